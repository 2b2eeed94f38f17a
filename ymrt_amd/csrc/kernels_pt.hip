// kernels_pt.hip -- wavefront path tracing on gfx950.
//
// Stages of one bounce pass (the reference's Ray_generation / Ray_trace /
// Ray_shading kernels, src/Shading.cu:1634-2302, with its inline shadow rays
// split into their own stage):
//   raygen  camera rays for a chunk of (pixel, sample) paths
//   extend  closest hit + homogeneous/grid free-flight sampling, miss -> env
//   shade   surface or volume vertex: NEE light and BSDF/phase samples written
//           as up to two shadow rays, next direction, Russian roulette
//   shadow  any-hit of each shadow ray (+ boundary-walk transmittance when
//           media are on), adds the NEE term in the reference's order
//   resolve per-pixel mean over the samples of the chunk (sample order fixed,
//           so frames are deterministic -- no float atomics)
// Path state is SoA in HBM indexed by the path's slot; queues hold slot ids
// and are compacted per wave64 with a ballot + one atomic per wave.  Every
// kernel fetches 64 items per wave from a device counter, so launches never
// need the host to know queue lengths (no host round trip per stage).
#include <mutex>
#include <unordered_map>

#include "dev_shade.hpp"
#include "dev_traverse.hpp"
#include "internal.hpp"
#include "pt.hpp"

namespace ym {

constexpr int kBlock = 64;

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Items a wave takes per work-fetch atomic: kSub sub-batches of 64.  Every
// global counter here is touched by all waves of the grid, and device-scope
// atomics on one address serialise across the chip, so each wave fetches
// kSub * 64 items at once and publishes each output queue's survivors of all
// sub-batches with one atomic.  Stages with little work per item take more
// per fetch (ray generation 32 sub-batches, free flight 8), where the one
// counter's atomics would otherwise set the pace; the default of 4 measured
// best for the others.
constexpr int kSub = 4;

template <int kS = kSub>
__device__ __forceinline__ uint32_t wave_fetch(uint32_t *work)
{
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(work, (uint32_t)(kBlock * kS));
    return __shfl(base, 0);
}

// Persistent loop over n items (one wave per block): body(i, p, out) handles
// item i < n, sets p (the path slot it appends) and out[q] for each output
// queue q.  Survivors of a fetch's sub-batches are staged in LDS and then
// published with one atomic per queue and a coalesced copy.
template <int kQueues, int kS = kSub, typename Body>
__device__ __forceinline__ void run_items(uint32_t *fetch, uint32_t n, uint32_t *const *counters,
                                          uint32_t *const *queues, Body &&body)
{
    constexpr int kQ = kQueues > 0 ? kQueues : 1;
    __shared__ uint32_t stage[kQ * kS * kBlock];
    const uint64_t below = (1ull << lane_id()) - 1ull;
    for (;;) {
        const uint32_t base = wave_fetch<kS>(fetch);
        if (base >= n) break;
        uint32_t filled[kQ] = {};
#pragma unroll 1
        for (int k = 0; k < kS; ++k) {
            const uint32_t first = base + (uint32_t)k * kBlock;
            if (first >= n) break;
            const uint32_t i = first + lane_id();
            uint32_t p = 0;
            bool out[kQ] = {};
            if (i < n) body(i, p, out);
            static_for<0, kQueues>([&](auto Q) {
                const uint64_t m = __ballot(out[Q()]);
                if (out[Q()]) stage[Q() * kS * kBlock + filled[Q()] + (uint32_t)__popcll(m & below)] = p;
                filled[Q()] += (uint32_t)__popcll(m);
            });
        }
        if constexpr (kQueues > 0) {
            __syncthreads();
            static_for<0, kQueues>([&](auto Q) {
                if (filled[Q()]) {
                    uint32_t dst = 0;
                    if (lane_id() == 0) dst = atomicAdd(counters[Q()], filled[Q()]);
                    dst = __shfl(dst, 0);
                    for (uint32_t j = lane_id(); j < filled[Q()]; j += kBlock)
                        queues[Q()][dst + j] = stage[Q() * kS * kBlock + j];
                }
            });
            __syncthreads();
        }
    }
}

// Lane refill: a kernel that keeps one traversal per lane lets finished lanes
// take new items while the others stay where they are in their trees, once
// kRefillLanes lanes are done.  Ray lengths vary by orders of magnitude, so a
// wave that waited for its slowest ray would run most of its steps with few
// lanes active.  refill() hands the wave's idle lanes the next items of its
// fetched range [cur, end), fetching kSub * 64 more as the range runs dry; it
// returns true on the lanes that took item *item.  Refilling at 32 (extend)
// and 16 (shadow) idle lanes measured fastest on the interior frame (extend:
// 16, 24, 28, 40 and 48 were 1-13 % slower; shadow: 8, 24 and 32 were
// slower or the same).
constexpr int kLaneIdle = 3;  // no ray (beside kTravContinue/Hit/Miss/Overflow)
template <int kRefillLanes>
struct RefillRange {
    uint32_t *work;
    uint32_t n;
    uint32_t cur, end;
    bool more;  // the work counter may hold items
    __device__ __forceinline__ bool refill(bool idle_lane, uint32_t &item)
    {
        const uint64_t below = (1ull << lane_id()) - 1ull;
        bool took = false;
        uint64_t idle = __ballot(idle_lane);
        while (idle && more) {
            if (cur >= end) {
                const uint32_t base = wave_fetch(work);
                if (base >= n) {
                    more = false;
                    break;
                }
                cur = base;
                end = base + (uint32_t)(kBlock * kSub) < n ? base + (uint32_t)(kBlock * kSub) : n;
            }
            const uint32_t take = min((uint32_t)__popcll(idle), end - cur);
            if (idle_lane && !took) {
                const uint32_t rank = (uint32_t)__popcll(idle & below);
                if (rank < take) {
                    item = cur + rank;
                    took = true;
                }
            }
            cur += take;
            idle = __ballot(idle_lane && !took);
        }
        return took;
    }
    // lanes to wait for before the next refill: all of them once the counter
    // is drained, since nothing could refill them
    __device__ __forceinline__ uint32_t want() const { return more ? (uint32_t)kRefillLanes : (uint32_t)kBlock; }
};

__device__ __forceinline__ void wave_add64(unsigned long long *counter, uint32_t v)
{
    // sum v over the wave, one atomic
    uint32_t s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0 && s) atomicAdd(counter, (unsigned long long)s);
}

// ------------------------------------------------------------ adjoint replay
// The replay pass of ymrt_render_adjoint runs the frame's paths a second time
// with the same samples (every random number is keyed by pixel, sample and
// vertex), so each path takes the same vertices and ps.out[p] already holds
// its radiance L_path from the first pass.  A gradient event whose factor
// multiplies every deposit after it adds adj . (L_path - L_now) * w; a term
// of one shadow ray's deposit D adds adj . D * w; adj is the pixel's
// dLoss/dImage over spp.  Sampling pdfs and MIS weights stay fixed.

// Pixel adjoint / spp of path slot p (64-bit slot arithmetic as k_raygen).
__device__ __forceinline__ f3 adj_pixel(const PtParams &prm, uint32_t p)
{
    const uint64_t gp = prm.chunk_base + p;
    const uint32_t spp = (uint32_t)prm.rs.spp;
    const uint64_t vpix = gp / spp;
    const uint32_t tile = (uint32_t)(vpix >> 8), within = (uint32_t)(vpix & 255u);
    const uint32_t ty = tile / prm.tiles_x, tx = tile - ty * prm.tiles_x;
    const uint32_t px = tx * 16u + (within & 15u), py = ty * 16u + (within >> 4);
    const float *a = prm.adj.img + ((size_t)py * prm.rs.width + px) * 3;
    const float inv = 1.0f / (float)spp;
    return f3{a[0] * inv, a[1] * inv, a[2] * inv};
}

// Gradient sums of one wave in a replay kernel: albedo rows and roughness
// entries go to LDS (kAdjLds floats when they fit) and reach global memory
// with one atomic per nonzero entry when the wave's work ends; texels and
// voxels go straight to global memory (their targets are spread out).
// kOn false: nothing is compiled.
template <bool kOn>
struct AdjSums {
    float *lds;
    const AdjParams *A;
    __device__ __forceinline__ void begin()
    {
        if constexpr (kOn) {
            if (A->lds) {
                for (uint32_t i = lane_id(); i < kAdjLds; i += kBlock) lds[i] = 0.0f;
                __syncthreads();
            }
        }
    }
    __device__ __forceinline__ void add(uint32_t i, float v)
    {
        if constexpr (kOn) {
            if (v == 0.0f) return;
            if (A->lds) atomicAdd(&lds[i], v);
            else if (i < A->n_alb * 3u) unsafeAtomicAdd(&A->g_alb[i], v);
            else unsafeAtomicAdd(&A->g_rough[i - A->n_alb * 3u], v);
        }
    }
    __device__ __forceinline__ void albedo(uint32_t row, int c, float v) { add(row * 3u + (uint32_t)c, v); }
    __device__ __forceinline__ void roughness(uint32_t entry, float v) { add(A->n_alb * 3u + entry, v); }
    __device__ __forceinline__ void end()
    {
        if constexpr (kOn) {
            if (A->lds) {
                __syncthreads();
                const uint32_t n = A->n_alb * 3u + A->n_rough;
                for (uint32_t i = lane_id(); i < n; i += kBlock) {
                    const float v = lds[i];
                    if (v == 0.0f) continue;
                    if (i < A->n_alb * 3u) unsafeAtomicAdd(&A->g_alb[i], v);
                    else unsafeAtomicAdd(&A->g_rough[i - A->n_alb * 3u], v);
                }
            }
        }
    }
};

// Density events of the trackers for the target volume's voxel grid: a real
// collision (probability sigma / sigma_m) adds scale * tsig / sigma, a null
// collision or a ratio-tracking step (1 - sigma / sigma_m) adds scale *
// -tsig / (sigma_m - sigma), tsig the target's sigma_t (gray); at most
// `left` events.
struct AdjTrackEvents {
    const DevScene &sc;
    float *g_vox;
    uint32_t vol;
    int grid;
    float tsig, scale;
    uint32_t left;
    int vox;
    __device__ __forceinline__ void note(int v, f3 po)
    {
        if ((uint32_t)v == vol) vox = grid_voxel(sc, grid, po);
    }
    __device__ __forceinline__ void real(float st, float sm)
    {
        (void)sm;
        if (vox >= 0 && st > 0.0f && left) {
            unsafeAtomicAdd(&g_vox[vox], scale * (tsig / st));
            --left;
        }
        vox = -1;
    }
    __device__ __forceinline__ void null(float st, float sm)
    {
        if (vox >= 0 && sm - st > 0.0f && left) {
            unsafeAtomicAdd(&g_vox[vox], scale * (-tsig / (sm - st)));
            --left;
        }
        vox = -1;
    }
};

__device__ __forceinline__ AdjTrackEvents adj_track(const DevScene &sc, const AdjParams &A, float scale,
                                                    uint32_t limit)
{
    const DevVolume &v = sc.volumes[A.vol];
    return AdjTrackEvents{sc, A.g_vox, A.vol, v.grid, v.sigma_t[0], scale, limit, -1};
}

// A shadow ray's transmittance records at most this many events; later ones
// are dropped (the CPU restatement of the adjoint drops the same).
constexpr uint32_t kAdjRayEvents = 1024;

// ------------------------------------------------------------ frame slots
// Multi-frame launches (ymrt_render_batch): one launch of a pass kernel serves
// every frame chunk of a round, each chunk a slot with its own scene,
// settings, counters and part of the path state.  A wave starts at its home
// slot -- the slot whose share of the launch's items holds the wave's index,
// so waves spread over the slots in proportion to their work -- drains it
// with the single-frame body, then visits the others in turn and helps with
// what is left there.  The slot index is wave-uniform, so the slot's fields
// come from scalar loads as kernel arguments do.
__device__ __forceinline__ uint32_t pass_counter(const FrameSlot &f, uint32_t pass, uint32_t k)
{
    return f.prm.counters[pass * kCtrStride + k];
}

// the work counter has passed the item count: nothing left to fetch
__device__ __forceinline__ bool work_drained(const uint32_t *work, uint32_t n)
{
    return __hip_atomic_load(work, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n;
}
__device__ __forceinline__ bool pass_drained(const FrameSlot &f, uint32_t pass, uint32_t work, uint32_t count)
{
    const uint32_t *c = f.prm.counters + pass * kCtrStride;
    return work_drained(c + work, c[count]);
}

template <typename Count, typename Drained, typename Body>
__device__ __forceinline__ void for_slots(const FrameSlot *__restrict__ fs, uint32_t nfs, Count count,
                                          Drained drained, Body body)
{
    uint32_t total = 0;
    for (uint32_t s = 0; s < nfs; ++s) total += count(fs[s]);
    if (total == 0) return;
    // waves beyond one per fetch of kSub * 64 items would find the stage
    // dry: they leave without touching its work counter (whose atomics
    // queue up behind each other), and the rest take every item
    if ((uint64_t)blockIdx.x * (uint64_t)(kBlock * kSub) >= (uint64_t)total) return;
    const uint32_t target = (uint32_t)(((uint64_t)blockIdx.x * total) / gridDim.x);
    uint32_t home = 0;
    for (uint32_t s = 0, acc = 0; s < nfs; ++s) {
        acc += count(fs[s]);
        if (acc > target) {
            home = s;
            break;
        }
    }
    for (uint32_t j = 0; j < nfs; ++j) {
        uint32_t s = home + j;
        s = __builtin_amdgcn_readfirstlane(s >= nfs ? s - nfs : s);
        const FrameSlot &f = fs[s];
        if (count(f) == 0 || drained(f)) continue;
        body(f);
    }
}

// ------------------------------------------------------------ state access
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__device__ __forceinline__ Transfer load_transfer(const PathState &ps, uint32_t p, uint32_t cap)
{
    Transfer t;
    const uint32_t h = ps.xhdr[p];
    t.count = (int)(h & 0xFFu);
    t.max_idx = (int)((h >> 8) & 0xFFu) - 1;
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        uint4 r = make_uint4(0u, 0u, 0u, 0u);
        if (i < t.count) r = ps.xrec[(size_t)i * cap + p];
        t.rec[i] = CameraRecord{r.x, __uint_as_float(r.y), r.z, (int32_t)r.w};
    }
    return t;
}
__device__ __forceinline__ void store_transfer(const PathState &ps, uint32_t p, uint32_t cap, const Transfer &t)
{
    ps.xhdr[p] = (uint32_t)t.count | ((uint32_t)(t.max_idx + 1) << 8);
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i)
        if (i < t.count)
            ps.xrec[(size_t)i * cap + p] = make_uint4(t.rec[i].prim, __float_as_uint(t.rec[i].ior),
                                                      t.rec[i].priority, (uint32_t)t.rec[i].vol);
}
__device__ __forceinline__ NearbyList load_nearby(const PathState &ps, uint32_t p, uint32_t cap, uint32_t hflags)
{
    NearbyList nb;
    nb.n = (int)((hflags >> 4) & 0xFu);
#pragma unroll
    for (int i = 0; i < kNearbyMax; ++i) {
        uint2 e = make_uint2(0u, 0u);
        if (i < nb.n) e = ps.nb[(size_t)i * cap + p];
        nb.prim[i] = e.x & 0x7FFFFFFFu;
        nb.back[i] = (int)(e.x >> 31);
        nb.t[i] = __uint_as_float(e.y);
    }
    return nb;
}
__device__ __forceinline__ Sampler load_sampler(const PathState &ps, uint32_t p)
{
    const uint4 s = ps.smp[p];
    Sampler r;
    r.index = (uint64_t)s.x | ((uint64_t)s.y << 32);
    r.dim = s.z;
    r.frame_type = s.w;
    return r;
}
__device__ __forceinline__ void store_sampler(const PathState &ps, uint32_t p, const Sampler &s)
{
    ps.smp[p] = make_uint4((uint32_t)s.index, (uint32_t)(s.index >> 32), s.dim, s.frame_type);
}

// info word
constexpr uint32_t kInfoNeverScatter = 1u << 8;
constexpr uint32_t kInfoPendingBunch = 1u << 10;
__device__ __forceinline__ int info_depth(uint32_t i) { return (int)(i & 0xFFu); }
__device__ __forceinline__ uint32_t info_vertex(uint32_t i) { return i >> 16; }

// Stage flag kind of a surface hit on a material of shading class cls.
__device__ __forceinline__ uint32_t shade_flag(uint32_t cls)
{
    return cls == kMatDiffuse ? kFlagDiffuse : cls == kMatMetal ? kFlagMetal : kFlagHit;
}

// hit flags
constexpr uint32_t kHitBack = 1u;
constexpr uint32_t kHitVolume = 2u;

// shadow flags: bits 0-3 shadow ray k recorded (slots 0-1 shape light,
// 2-3 distant light), 4-7 ray k crosses the surface, 8-11 ray k reached the
// light past solid geometry, 12 volume vertex, 13 path ends after this
// shadow stage, 16-23 vertex depth
constexpr uint32_t kShActiveMask = 0xFu;
constexpr uint32_t kShTransmit0 = 1u << 4;
constexpr uint32_t kShClear0 = 1u << 8;
constexpr uint32_t kShClearMask = 0xFu << 8;
constexpr uint32_t kShVolume = 1u << 12;
constexpr uint32_t kShTerminate = 1u << 13;
constexpr uint32_t kShDepthShift = 16;

// indirect_light_clamp (src/Shading.cu:61-81).  Its powers use the GPU's
// float powf: the double-precision form (ym_powf) would raise the registers
// of every kernel that deposits radiance, for a setting that is off in the
// measured workloads.
__device__ __forceinline__ f3 light_clamp(const ymrt_render_settings &rs, f3 c, int depth)
{
    if (!(rs.enable_indirect_clamp && depth > 1)) return c;
    if (!(rs.indirect_limit > 0.0f)) return s3(0.0f);
    const float lum = luminance(c);
    if (lum <= 1E-10f) return c;
    const float lim = rs.indirect_limit;
    const float pw = gmin(0.05f + 0.25f * powf(lum / lim, 1.5f), 1.0f);
    const float sf = powf(lim / (lim + lum), pw);
    return c * sf;
}

// ------------------------------------------------------------ raygen
__device__ __forceinline__ void raygen_body(const DevScene &sc, const PtParams &prm, const PathState &ps)
{
    uint32_t *cnt = prm.counters;  // pass 0 slots
    uint32_t *const ctrs[1] = {cnt + kCtrActive};
    uint32_t *const qs[1] = {ps.queue[0]};
    run_items<1, 32>(cnt + kCtrWorkRaygen, prm.chunk_paths, ctrs, qs, [&](uint32_t p, uint32_t &slot, bool(&out)[1]) {
        slot = p;
        ps.flag[p] = 0;
        {
            const uint64_t gp = prm.chunk_base + p;
            const uint32_t spp = (uint32_t)prm.rs.spp;
            const uint64_t vpix = gp / spp;
            const uint32_t s = (uint32_t)(gp - vpix * spp);
            const uint32_t tile = (uint32_t)(vpix >> 8), within = (uint32_t)(vpix & 255u);
            const uint32_t ty = tile / prm.tiles_x, tx = tile - ty * prm.tiles_x;
            const uint32_t px = tx * 16u + (within & 15u), py = ty * 16u + (within >> 4);
            if (px < prm.rs.width && py < prm.rs.height) {
                out[0] = true;
                const uint32_t pixel_index = py * prm.rs.width + px;
                Sampler smp;
                float ox, oy;
                if (prm.rs.sampler_type == YMRT_SAMPLER_HALTON) {
                    smp.index = prm.henum.index(px, py, (uint32_t)((prm.frame - 1) * prm.rs.spp) + s);
                    smp.dim = 2;
                    smp.frame_type = kSamplerHalton;
                    ox = radical_inverse_plain<2>(smp.index);
                    oy = radical_inverse_plain<3>(smp.index);
                    ox = __builtin_fmaf(ox, prm.henum.x_scale, -(float)px);
                    oy = __builtin_fmaf(oy, prm.henum.y_scale, -(float)py);
                    ox = gclamp(ox, 0.000001f, 0.999999f);
                    oy = gclamp(oy, 0.000001f, 0.999999f);
                } else {
                    smp.index = (uint64_t)pixel_index | ((uint64_t)s << 32);
                    smp.dim = 0;
                    smp.frame_type = (uint32_t)(prm.frame - 1) << 1;
                    ox = sampler_next(sc, smp);
                    oy = sampler_next(sc, smp);
                }
                const DevCamera &cam = sc.camera;
                const float nx = ((float)px + ox) / (float)prm.rs.width;
                const float ny = ((float)py + oy) / (float)prm.rs.height;
                const float height = 2.0f * cam.tan_half_fov;
                const float width = height * cam.aspect;
                const f3 dir = normalize(-cam.w + ((nx - 0.5f) * width) * cam.u + ((ny - 0.5f) * height) * cam.v);
                ps.ray_o[p] = f4(cam.position, 0.0f);
                ps.ray_d[p] = f4(dir, 0.0f);
                ps.last[p] = make_uint2(kEmpty, kEmpty);
                ps.beta[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                ps.L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                store_sampler(ps, p, smp);
                ps.info[p] = 1u | kInfoNeverScatter;
                Transfer t;
                t.count = cam.record_count;
                t.max_idx = cam.max_idx;
                for (int i = 0; i < kMaxBoundaryRecord; ++i) t.rec[i] = cam.rec[i];
                store_transfer(ps, p, prm.cap, t);
            }
        }
    });
}


__global__ __launch_bounds__(kBlock) void k_raygen_m(const FrameSlot *__restrict__ fs, uint32_t nfs)
{
    for_slots(
        fs, nfs, [](const FrameSlot &f) { return f.prm.chunk_paths; },
        [](const FrameSlot &f) { return work_drained(f.prm.counters + kCtrWorkRaygen, f.prm.chunk_paths); },
        [](const FrameSlot &f) { raygen_body(f.sc, f.prm, f.ps); });
}



// ------------------------------------------------------------ extend
// Closest hit for every queued path (Ray_trace_path_tracing,
// src/Shading.cu:1765-1901).  With kMedia, paths that carry medium records
// hand their hit or miss to k_medium, which samples free flight; the others
// resolve misses here and queue surface hits for shading.  Visited wide nodes
// and tested triangles are summed per wave into the frame statistics.
// Each lane keeps its own traversal and is refilled as it finishes
// (RefillRange).

template <bool kNearby, bool kMedia, bool kSphere, int kStack>
__device__ __forceinline__ void extend_body(const DevScene &sc, const PtParams &prm, const PathState &ps, uint32_t pass)
{
    __shared__ uint32_t stack[kStackWindow * kBlock];
    __shared__ uint32_t nbp[(kNearby ? kNearbyCap : 1) * kBlock];
    __shared__ float nbt[(kNearby ? kNearbyCap : 1) * kBlock];
    __shared__ int nbb[(kNearby ? kNearbyCap : 1) * kBlock];
    __shared__ float pre_park[kPreSlots * kBlock];
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    const uint32_t *qin = ps.queue[pass & 1];
    LaneScratch<kBlock, kStackWindow, true> ls{stack, nbp, nbt, nbb, (int)lane_id(), ps.spill, ps.spill_threads,
                                               blockIdx.x * kBlock + threadIdx.x, pre_park};
    const ymrt_render_settings &rs = prm.rs;
    uint32_t rays = 0, cam_rays = 0;
    TraceCounters tc{0, 0};
    RefillRange<32> rr{cnt + kCtrWorkExtend, cnt[kCtrActive], 0u, 0u, true};

    Traversal<kModeAll, false, kNearby, true, kSphere, kStack> tv;
    uint32_t p = 0;
    int r = kLaneIdle;
    for (;;) {
        uint32_t item = 0;
        if (rr.more && rr.refill(r == kLaneIdle, item)) {
            p = gload(qin + item);
            const float4 o4 = gload(ps.ray_o + p), d4 = gload(ps.ray_d + p);
            const uint2 last = gload(ps.last + p);
            Query q;
            q.o = xyz(o4);
            q.d = xyz(d4);
            q.tmin = kTMin;
            q.tmax = kTMax;
            q.last_inst = last.x;
            q.last_tri = last.y;
            tv.hit.back = 0;
            tv.hit.inst = kEmpty;
            tv.hit.tri = 0;
            tv.hit.b0 = tv.hit.b1 = tv.hit.b2 = 0.0f;
            tv.nb_count = 0;
            r = tv.init(sc, q);
        }
        const uint32_t want = rr.want();
        for (;;) {
            const uint64_t busy = __ballot(r == kTravContinue);
            if (busy == 0 || (uint32_t)(kBlock - __popcll(busy)) >= want) break;
            wave_step<16, true>(tv, r, sc, ls, tc);
        }
        if (r != kTravContinue && r != kLaneIdle) {
            if (r == kTravOverflow) atomicOr(prm.error, 1u);
            const bool hit = r == kTravHit;
            r = kLaneIdle;
            uint32_t info = gload(ps.info + p);
            if (info & kInfoPendingBunch) {
                // the boundary bunch of the previous vertex reads that
                // vertex's nearby list, which the writes below replace
                Transfer t = load_transfer(ps, p, prm.cap);
                const NearbyList nb = load_nearby(ps, p, prm.cap, ps.hit_a[p].w);
                boundary_bunch(sc, t, nb);
                store_transfer(ps, p, prm.cap, t);
                info &= ~kInfoPendingBunch;
                ps.info[p] = info;
            }
            rays += 1;
            cam_rays += (info_depth(info) == 1 && (info & kInfoNeverScatter)) ? 1u : 0u;
            const Hit &h = tv.hit;
            if (pass == 0 && prm.aov) ps.first_hit[p] = hit ? h.inst : kEmpty;
            const bool in_medium = kMedia && prm.media && (gload(ps.xhdr + p) & 0xFFu) != 0;
            if (!hit && !in_medium) {
                f3 Lv = xyz(gload(ps.L + p));
                if (rs.enable_env_light)
                    Lv = Lv + light_clamp(rs, xyz(gload(ps.beta + p)) * background(tv.q.d), info_depth(info));
                gstore(ps.out + p, f4(Lv, 0.0f));
            } else {
                // a path in a medium may have no surface hit (h.inst empty);
                // one outside a medium always has one here
                const uint32_t mat = hit ? gload(&sc.instances[h.inst].material) : kEmpty;
                gstore(ps.flag + p, (pass << kFlagBits) | (in_medium ? kFlagMedium : shade_flag(gload(sc.mat_class + mat))));
                uint32_t flags = h.back ? kHitBack : 0u;
                bool listed = false;
                const int nbc = tv.nb_count;
                if constexpr (kNearby) {
                    if (hit) {
                        listed = true;
                        flags |= (uint32_t)nbc << 4;
                        static_for<0, kNearbyCap>([&](auto K) {
                            if (K() < nbc)
                                gstore(ps.nb + (size_t)K() * prm.cap + p,
                                       make_uint2(nbp[K() * kBlock + lane_id()] |
                                                      ((uint32_t)(nbb[K() * kBlock + lane_id()] != 0) << 31),
                                                  __float_as_uint(nbt[K() * kBlock + lane_id()])));
                        });
                    }
                }
                if (!listed && hit) {
                    // without the list the closest instance is its own nearby hit
                    flags |= 1u << 4;
                    gstore(ps.nb + p, make_uint2(h.inst | ((uint32_t)(h.back != 0) << 31), __float_as_uint(h.t)));
                }
                gstore(ps.hit_a + p, make_uint4(__float_as_uint(hit ? h.t : kTMax), hit ? h.inst : kEmpty, h.tri, flags));
                gstore(ps.hit_b + p, make_float4(h.b0, h.b1, h.b2, 0.0f));
            }
        }
        if (!rr.more && __ballot(r != kLaneIdle) == 0) break;
    }
    wave_add64(prm.stats + kStatExtension, rays);
    wave_add64(prm.stats + kStatCamera, cam_rays);
    wave_add64(prm.stats + kStatExtendNodes, tc.nodes);
    wave_add64(prm.stats + kStatExtendTris, tc.tris);
}


template <bool kNearby, bool kMedia, bool kSphere, int kStack>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kSphere ? 1 : 4))) void k_extend_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass)
{
    for_slots(
        fs, nfs, [&](const FrameSlot &f) { return pass_counter(f, pass, kCtrActive); },
        [&](const FrameSlot &f) { return pass_drained(f, pass, kCtrWorkExtend, kCtrActive); },
        [&](const FrameSlot &f) { extend_body<kNearby, kMedia, kSphere, kStack>(f.sc, f.prm, f.ps, pass); });
}



// ------------------------------------------------------------ compaction
// Gathers the paths flagged for the next stage of this pass (kShade: surface
// and volume shading; otherwise free flight) in slot order: the traversal
// kernels finish paths in no fixed order, and read in slot order the next
// stage's lanes (and neighbouring waves) load neighbouring path-state
// entries together.  A scan of every slot's flag (4 B) costs less than the
// scattered reads it saves.  No atomics: k_compact_count sums the matches of
// each block of kCompactSpan slots, k_compact_scatter writes a block's
// matches at the sum of the counts before it, and the block holding the last
// slot sets the queue lengths.
// kMode: kCompactMedium, kCompactShade (four queues) or kCompactTransmit
enum : int { kCompactMedium = 0, kCompactShade = 1, kCompactTransmit = 2 };
template <int kMode>
struct Compact {
    static constexpr int kQ = kMode == kCompactShade ? 4 : 1;
    __device__ static void match(uint32_t f, uint32_t tag, bool (&m)[kQ])
    {
        if constexpr (kMode == kCompactShade) {
            m[0] = f == (tag | kFlagHit);
            m[1] = f == (tag | kFlagVolume);
            m[2] = f == (tag | kFlagDiffuse);
            m[3] = f == (tag | kFlagMetal);
        } else {
            m[0] = f == (tag | (kMode == kCompactMedium ? kFlagMedium : kFlagTransmit));
        }
    }
};

template <int kMode>
__device__ __forceinline__ void compact_count_body(const PtParams &prm, const PathState &ps, uint32_t pass, uint32_t b)
{
    constexpr int kQ = Compact<kMode>::kQ;
    constexpr uint32_t kWaves = kCompactThreads / kBlock;
    __shared__ uint32_t red[kQ][kWaves];
    const uint32_t n = prm.chunk_paths;
    const uint32_t first = b * kCompactSpan;
    if (first >= n) return;
    const uint32_t last = min(first + kCompactSpan, n);
    const uint32_t tag = pass << kFlagBits;
    uint32_t c[kQ] = {};
    for (uint32_t i = first + threadIdx.x; i < last; i += kCompactThreads) {
        bool m[kQ];
        Compact<kMode>::match(ps.flag[i], tag, m);
        static_for<0, kQ>([&](auto Q) { c[Q()] += m[Q()] ? 1u : 0u; });
    }
    static_for<0, kQ>([&](auto Q) {
        uint32_t v = c[Q()];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane_id() == 0) red[Q()][threadIdx.x / kBlock] = v;
    });
    __syncthreads();
    if (threadIdx.x < (uint32_t)kQ) {
        uint32_t v = 0;
        for (uint32_t w = 0; w < kWaves; ++w) v += red[threadIdx.x][w];
        ps.ccount[threadIdx.x * prm.ccount_stride + b] = v;
    }
}


// one row of the grid (blockIdx.y) per slot
template <int kMode>
__global__ __launch_bounds__(kCompactThreads) void k_compact_count_m(const FrameSlot *__restrict__ fs, uint32_t pass)
{
    const FrameSlot &f = fs[blockIdx.y];
    compact_count_body<kMode>(f.prm, f.ps, pass, blockIdx.x);
}

template <int kMode>
__device__ __forceinline__ void compact_scatter_body(const PtParams &prm, const PathState &ps, uint32_t pass,
                                                     uint32_t b)
{
    constexpr int kQ = Compact<kMode>::kQ;
    constexpr uint32_t kWaves = kCompactThreads / kBlock;
    __shared__ uint32_t red[kQ][kWaves];
    __shared__ uint32_t base[kQ];
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    const uint32_t n = prm.chunk_paths;
    const uint32_t first = b * kCompactSpan;
    if (first >= n) return;
    const uint32_t last = min(first + kCompactSpan, n);
    const uint32_t stride = prm.ccount_stride;
    uint32_t *const qs[4] = {kMode == kCompactShade    ? ps.q_hit
                             : kMode == kCompactMedium ? ps.q_medium
                                                       : ps.q_transmit,
                             ps.q_hit_vol, ps.q_hit_diffuse, ps.q_hit_metal};
    const uint32_t ctr[4] = {kMode == kCompactShade ? kCtrHit : kMode == kCompactMedium ? kCtrMedium : kCtrTransmit,
                             kCtrHitVol, kCtrHitDiffuse, kCtrHitMetal};
    const uint32_t tag = pass << kFlagBits;
    const uint32_t wave = threadIdx.x / kBlock;
    const uint64_t below = (1ull << lane_id()) - 1ull;
    // this block's offsets: the counts of the blocks before it, summed by
    // k_compact_scan
    if (threadIdx.x < (uint32_t)kQ) base[threadIdx.x] = ps.ccount[threadIdx.x * stride + b];
    __syncthreads();
    uint32_t off[kQ];
    static_for<0, kQ>([&](auto Q) { off[Q()] = base[Q()]; });
    for (uint32_t t0 = first; t0 < last; t0 += kCompactThreads) {
        const uint32_t i = t0 + threadIdx.x;
        uint32_t p = 0;
        bool m[kQ] = {};
        if (i < last) {
            p = i;
            Compact<kMode>::match(ps.flag[p], tag, m);
        }
        uint32_t rank[kQ];
        __syncthreads();  // red[] of the previous tile has been read
        static_for<0, kQ>([&](auto Q) {
            const uint64_t mask = __ballot(m[Q()]);
            rank[Q()] = (uint32_t)__popcll(mask & below);
            if (lane_id() == 0) red[Q()][wave] = (uint32_t)__popcll(mask);
        });
        __syncthreads();
        static_for<0, kQ>([&](auto Q) {
            uint32_t before = 0, total = 0;
            for (uint32_t w = 0; w < kWaves; ++w) {
                const uint32_t v = red[Q()][w];
                before += w < wave ? v : 0u;
                total += v;
            }
            if (m[Q()]) qs[Q()][off[Q()] + before + rank[Q()]] = p;
            off[Q()] += total;
        });
    }
    if (last == n && threadIdx.x < (uint32_t)kQ) cnt[ctr[threadIdx.x]] = off[threadIdx.x];
}


// The per-block counts of one slot (blockIdx.y; one block per queue) turned
// into their exclusive prefix sums in place, so that each scatter block reads
// its offsets instead of summing the counts of all blocks before it (a sum
// whose total work grows with the square of the block count).
constexpr uint32_t kScanThreads = 1024;
template <int kMode>
__global__ __launch_bounds__(kScanThreads) void k_compact_scan_m(const FrameSlot *__restrict__ fs)
{
    const FrameSlot &f = fs[blockIdx.y];
    constexpr int kQ = Compact<kMode>::kQ;
    constexpr uint32_t kWaves = kScanThreads / kBlock;
    __shared__ uint32_t wsum[kWaves];
    const uint32_t nb = (f.prm.chunk_paths + kCompactSpan - 1) / kCompactSpan;
    const uint32_t stride = f.prm.ccount_stride;
    const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
    const uint32_t lo = min(threadIdx.x * per, nb), hi = min(lo + per, nb);
    const uint32_t wave = threadIdx.x / kBlock;
    {
        const uint32_t q = blockIdx.x;  // one block per queue
        if (q >= (uint32_t)kQ) return;
        uint32_t *c = f.ps.ccount + (size_t)q * stride;
        uint32_t own = 0;
        for (uint32_t j = lo; j < hi; ++j) own += c[j];
        uint32_t incl = own;
        for (int off = 1; off < kBlock; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if ((int)lane_id() >= off) incl += y;
        }
        if (lane_id() == kBlock - 1) wsum[wave] = incl;
        __syncthreads();
        uint32_t before = incl - own;
        for (uint32_t w = 0; w < wave; ++w) before += wsum[w];
        for (uint32_t j = lo; j < hi; ++j) {
            const uint32_t v = c[j];
            c[j] = before;
            before += v;
        }
    }
}

// one row of the grid (blockIdx.y) per slot
template <int kMode>
__global__ __launch_bounds__(kCompactThreads) void k_compact_scatter_m(const FrameSlot *__restrict__ fs, uint32_t pass)
{
    const FrameSlot &f = fs[blockIdx.y];
    compact_scatter_body<kMode>(f.prm, f.ps, pass, blockIdx.x);
}

// ------------------------------------------------------------ medium
// Free flight through the media of the path's top-priority records
// (SampleVolumeScattering, src/VolumeEval.cuh:217-251, called from
// src/Shading.cu:1800-1840): scattering vertices go to the volume shading
// queue, surviving rays to surface shading or, on a miss, to the environment.
// kHet: 1 the scene has voxel-grid media (delta tracking), 2 texture-density
// media as well (the texture graph evaluator and its LDS stack are compiled
// in); 0 every medium is homogeneous and the tracker is compiled out.
// kAdj (the adjoint's replay): the delta tracker's collisions in the target
// grid volume are density gradient events.  The homogeneous variant is held
// to 5 waves per SIMD (96 VGPRs, no spills): it waits on memory most of the
// time, and the fifth wave measured 1.5 % faster than four.
template <int kHet, bool kAdj = false>
__device__ __forceinline__ void medium_body(const DevScene &sc, const PtParams &prm, const PathState &ps, uint32_t pass)
{
    constexpr bool kGrid = kHet > 0;
    __shared__ float tex_stk[kHet == 2 ? kTexStack * 3 * kBlock : 1];
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    const uint32_t n = cnt[kCtrMedium];
    const ymrt_render_settings &rs = prm.rs;
    run_items<0, 8>(cnt + kCtrWorkMedium, n, nullptr, nullptr, [&](uint32_t i, uint32_t &p, bool(&)[1]) {
        {
            p = ps.q_medium[i];
            const uint4 ha = ps.hit_a[p];
            const bool hit = ha.y != kEmpty;
            const f3 o = xyz(ps.ray_o[p]), d = xyz(ps.ray_d[p]);
            f3 beta = xyz(ps.beta[p]);
            bool vol_hit = false;
            float dist = 0.0f;
            const Transfer t = load_transfer(ps, p, prm.cap);
            int vi[kMaxBoundaryRecord];
            const int nv = t.volumes(vi, sc.volume_count);
            if (nv > 0) {
                Sampler smp = load_sampler(ps, p);
                bool het = false;
                if constexpr (kGrid) {
                    static_for<0, kMaxBoundaryRecord>([&](auto K) {
                        if (K() < nv && vol_is_het(sc, vi[K()])) het = true;
                    });
                }
                f3 trw;
                const float tmax = hit ? __uint_as_float(ha.x) : kTMax;
                if constexpr (kGrid && kAdj) {
                    // every later deposit of the path depends on this flight
                    float scale = 0.0f;
                    if (het && prm.adj.vol != kEmpty) {
                        const f3 dl = adj_pixel(prm, p) * (xyz(ps.out[p]) - xyz(ps.L[p]));
                        scale = dl.x + dl.y + dl.z;
                    }
                    if (scale != 0.0f) {
                        AdjTrackEvents ev = adj_track(sc, prm.adj, scale, 0xFFFFFFFFu);
                        vol_hit = sample_heterogeneous<kHet == 2>(sc, nv, vi, o, d, tmax, trw, dist, smp, tex_stk, ev);
                    } else {
                        vol_hit = het ? sample_heterogeneous<kHet == 2>(sc, nv, vi, o, d, tmax, trw, dist, smp, tex_stk)
                                      : sample_homogeneous(sc, nv, vi, tmax, trw, dist, smp);
                    }
                } else if constexpr (kGrid) {
                    vol_hit = het ? sample_heterogeneous<kHet == 2>(sc, nv, vi, o, d, tmax, trw, dist, smp, tex_stk)
                                  : sample_homogeneous(sc, nv, vi, tmax, trw, dist, smp);
                } else {
                    vol_hit = sample_homogeneous(sc, nv, vi, tmax, trw, dist, smp);
                }
                beta = beta * trw;
                ps.beta[p] = f4(beta, 0.0f);
                store_sampler(ps, p, smp);
            }
            if (vol_hit) {
                ps.flag[p] = (pass << kFlagBits) | kFlagVolume;
                const f3 vpos = o + d * dist;
                ps.hit_a[p] = make_uint4(ha.x, ha.y, ha.z, ha.w | kHitVolume);
                ps.hit_b[p] = make_float4(vpos.x, vpos.y, vpos.z, 0.0f);
            } else if (hit) {
                ps.flag[p] = (pass << kFlagBits) | shade_flag(sc.mat_class[sc.instances[ha.y].material]);
            } else {
                f3 Lv = xyz(ps.L[p]);
                if (rs.enable_env_light) Lv = Lv + light_clamp(rs, beta * background(d), info_depth(ps.info[p]));
                ps.out[p] = f4(Lv, 0.0f);
            }
        }
    });
}


template <int kHet, bool kAdj = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kHet ? 1 : 5))) void k_medium_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass)
{
    for_slots(
        fs, nfs, [&](const FrameSlot &f) { return pass_counter(f, pass, kCtrMedium); },
        [&](const FrameSlot &f) { return pass_drained(f, pass, kCtrWorkMedium, kCtrMedium); },
        [&](const FrameSlot &f) { medium_body<kHet, kAdj>(f.sc, f.prm, f.ps, pass); });
}



// ------------------------------------------------------------ shade
// Adjoint replay, a surface vertex whose diffuse albedo is the target image
// texture (every material using it is Lambert-only): the albedo is a bilinear
// blend of four texels, so each tap's texel channel c gets dl_c * w / albedo_c
// (a one-channel texture: the sum over the colour channels).
__device__ __forceinline__ void adj_texels(const DevScene &sc, const AdjParams &A, f2 uv, f3 dl)
{
    const DevTexture &t = sc.textures[A.tex];
    const f3 al = image_eval(sc, t, uv);
    uint32_t off[4];
    float w[4];
    image_taps(t, uv, off, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!(w[k] > 0.0f)) continue;
        if (t.channels == 1) {
            const float v = (al.x > 0.0f ? dl.x * (w[k] / al.x) : 0.0f) + (al.y > 0.0f ? dl.y * (w[k] / al.y) : 0.0f) +
                            (al.z > 0.0f ? dl.z * (w[k] / al.z) : 0.0f);
            if (v != 0.0f) unsafeAtomicAdd(&A.g_tex[off[k]], v);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float a = comp(al, c);
                if (!(a > 0.0f)) continue;
                const float v = comp(dl, c) * (w[k] / a);
                if (v != 0.0f) unsafeAtomicAdd(&A.g_tex[off[k] + (uint32_t)c], v);
            }
        }
    }
}

// Records shadow ray k of path p (origin, direction, maximum distance, BSDF or
// phase value, clamped Li, MIS weight) straight to its slot in HBM.  The
// direction is renormalised as the reference's Ray constructor does
// (src/Ray.h:28-29): a light-sample direction is already unit length up to
// rounding, and those last bits decide self-hits of rays that leave a surface
// with no spawn offset.
struct ShadowEmit {
    const PathState &ps;
    uint32_t cap, p;
    uint32_t flags;
    // g: the adjoint replay's d ln f / d r of the ray's BSDF value (0 otherwise)
    __device__ __forceinline__ void emit(int k, f3 orig, f3 dir, float tmax, bool transmit, f3 f, f3 li, float mis,
                                         float g = 0.0f)
    {
        const size_t o = (size_t)k * cap + p;
        ps.sh_o[o] = f4(orig, tmax);
        ps.sh_d[o] = f4(normalize(dir), 0.0f);
        ps.sh_f[o] = f4(f, mis);
        ps.sh_li[o] = f4(li, g);
        flags |= (1u << k) | (transmit ? (kShTransmit0 << k) : 0u);
    }
};

// Distance to a shape-light sample, shortened by SHADOW_RAY_CLAMP
// (src/Shading.cu:337).
__device__ __forceinline__ float shadow_tmax(f3 pos, f3 ldir, f3 lpos, f3 lerr, f3 lng)
{
    return length(offset_ray_origin(lpos, lerr, -ldir, lng) - pos) * kShadowRayClamp;
}

// kVolume: volume scattering vertices (Ray_shading_path_tracing's volume branch,
// src/Shading.cu:2070-2130, with EvalVolumeDistantLight 202-290 and
// EvalVolumeShapeLight 419-513); otherwise surface vertices (2132-2302 with
// EvalDistantLight 83-198 and EvalShapeLight 294-417).
// kGraph: the scene has procedural texture nodes or normal / bump maps (the
// variant without them keeps the image-texture fast path only).  kClass:
// kMatDiffuse or kMatMetal shades the hits on materials of that class (their
// own queue) with the lobes the class leaves fixed at compile time, which
// gives the general variant's floats with fewer live registers.  The
// variants without the graph evaluator are held to 168 VGPRs (3 waves per
// SIMD, a few spilled registers): the shading stage waits on dependent
// loads, and the third wave hides more of them than the spills cost (6 %
// faster on the interior frame; 4 waves spill too much).
// kAdj: the adjoint's replay pass (surface vertices): albedo, texel and
// roughness gradient events.
template <bool kVolume, bool kGraph, uint32_t kClass = kMatGeneral, bool kAdj = false>
__device__ __forceinline__ void shade_body(const DevScene &sc, const PtParams &prm, const PathState &ps, uint32_t pass)
{
    __shared__ float tex_stk[(kVolume || !kGraph) ? 1 : kTexStack * 3 * kBlock];
    __shared__ float adj_lds[kAdj ? kAdjLds : 1];
    AdjSums<kAdj> sums{adj_lds, &prm.adj};
    sums.begin();
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    uint32_t *cnt_next = prm.counters + (pass + 1) * kCtrStride;
    constexpr bool kDiffuse = kClass == kMatDiffuse, kMetal = kClass == kMatMetal;
    const uint32_t n = cnt[kVolume ? kCtrHitVol : kDiffuse ? kCtrHitDiffuse : kMetal ? kCtrHitMetal : kCtrHit];
    const uint32_t *q = kVolume ? ps.q_hit_vol : kDiffuse ? ps.q_hit_diffuse : kMetal ? ps.q_hit_metal : ps.q_hit;
    const ymrt_render_settings &rs = prm.rs;
    const uint32_t cap = prm.cap;
    uint32_t *const ctrs[2] = {cnt + kCtrShadow, cnt_next + kCtrActive};
    uint32_t *const qs[2] = {ps.q_shadow, ps.queue[(pass + 1) & 1]};
    constexpr uint32_t kLobes = kDiffuse ? 1u << kSlotDiffuse
                                : kMetal ? (1u << kSlotMetal) | (1u << kSlotDiffuse)
                                         : 0u;
    run_items<2>(cnt + (kVolume     ? kCtrWorkShadeVol
                        : kDiffuse ? kCtrWorkShadeDiffuse
                        : kMetal   ? kCtrWorkShadeMetal
                                   : kCtrWorkShade),
                 n, ctrs, qs,
                 [&](uint32_t i, uint32_t &p, bool(&out)[2]) {
        {
            p = q[i];
            uint32_t info = ps.info[p];
            int depth = info_depth(info);
            bool never_scatter = (info & kInfoNeverScatter) != 0;
            const uint32_t vertex = info_vertex(info) + 1;
            const uint4 ha = ps.hit_a[p];
            const float4 hb = ps.hit_b[p];
            const f3 ray_d = xyz(ps.ray_d[p]);
            f3 beta = xyz(ps.beta[p]);
            f3 L = xyz(ps.L[p]);
            Sampler smp = load_sampler(ps, p);
            ShadowEmit so{ps, cap, p, (uint32_t)depth << kShDepthShift};
            bool alive = false;       // a next extension ray
            const f3 sh_beta = beta;
            float sel_pdf = 0.0f;
            f3 next_o = s3(0.0f), next_d = s3(0.0f);
            uint32_t next_li = kEmpty, next_lt = kEmpty;
            bool pending = false;
            if constexpr (kVolume) {
                const Transfer t = load_transfer(ps, p, cap);
                int vi[kMaxBoundaryRecord];
                float gs[kMaxBoundaryRecord], vw[kMaxBoundaryRecord];
                const int nv = t.volumes(vi, sc.volume_count);
#pragma unroll
                for (int k = 0; k < kMaxBoundaryRecord; ++k) {
                    gs[k] = k < nv ? sc.volumes[vi[k]].g : 0.0f;
                    vw[k] = k < nv ? 1.0f / (float)nv : 0.0f;
                }
                const f3 vpos = xyz(hb);
                so.flags |= kShVolume;
                if (rs.enable_distant_light && sc.distant_count > 0 && nv > 0) {
                    const DevDistant &dl = sc.distant[pick_distant(sc.distant_count, sampler_next(sc, smp))];
                    {
                        f3 ldir = s3(0.0f), Li = s3(0.0f);
                        float lpdf = 0.0f;
                        const float u0 = sampler_next(sc, smp);
                        const float u1 = sampler_next(sc, smp);
                        distant_sample_li(dl, u0, u1, ldir, Li, lpdf);
                        float wpdf = 0.0f;
                        const float ph = mixed_phase_eval(vw, gs, nv, -ray_d, ldir, wpdf);
                        if (ph > kMinColorEpsilon)
                            so.emit(2, vpos, ldir, kTMax, false, s3(gmin(ph, kMaxColorClamp)),
                                    min3(Li, s3(kMaxColorClamp)), power_heuristic(lpdf, wpdf));
                    }
                    {
                        f3 pwi = s3(0.0f);
                        float ppdf = 0.0f, pw = 0.0f;
                        const float u0 = sampler_next(sc, smp);
                        const float u1 = sampler_next(sc, smp);
                        if (mixed_phase_sample(vw, gs, nv, u0, u1, -ray_d, pw, pwi, ppdf)) {
                            float lpdf = 0.0f;
                            const f3 Li = distant_eval_li(dl, pwi, lpdf);
                            if (maxcomp(Li) > kMinColorEpsilon)
                                so.emit(3, vpos, pwi, kTMax, false, s3(gmin(pw, kMaxColorClamp)),
                                        min3(Li, s3(kMaxColorClamp)), power_heuristic(ppdf, lpdf));
                        }
                    }
                }
                if (rs.enable_shape_light && sc.light_count > 0 && nv > 0) {
                    const int li = sample_light_power(sc, sampler_next(sc, smp), sel_pdf);
                    const DevLightRec &light = sc.light_recs[li];
                    {
                        f3 ldir = s3(0.0f), lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f), Li = s3(0.0f);
                        float lpdf = 0.0f;
                        const float u0 = sampler_next(sc, smp);
                        const float u1 = sampler_next(sc, smp);
                        const bool valid = light_sample_li(light, vpos, u0, u1, ldir, lpos, lerr, lng, Li, lpdf);
                        float wpdf = 0.0f;
                        const float ph = mixed_phase_eval(vw, gs, nv, -ray_d, ldir, wpdf);
                        if (ph > kMinColorEpsilon && valid)
                            so.emit(0, vpos, ldir, shadow_tmax(vpos, ldir, lpos, lerr, lng), false,
                                    s3(gmin(ph, kMaxColorClamp)), min3(Li, s3(kMaxColorClamp)),
                                    power_heuristic(lpdf, wpdf));
                    }
                    {
                        f3 pwi = s3(0.0f);
                        float ppdf = 0.0f, pw = 0.0f;
                        const float u0 = sampler_next(sc, smp);
                        const float u1 = sampler_next(sc, smp);
                        if (mixed_phase_sample(vw, gs, nv, u0, u1, -ray_d, pw, pwi, ppdf)) {
                            f3 lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f);
                            float lpdf = 0.0f;
                            const f3 Li = light_eval_li(light, vpos, pwi, lpos, lerr, lng, lpdf);
                            if (maxcomp(Li) > kMinColorEpsilon)
                                so.emit(1, vpos, pwi, shadow_tmax(vpos, pwi, lpos, lerr, lng), false,
                                        s3(gmin(pw, kMaxColorClamp)), min3(Li, s3(kMaxColorClamp)),
                                        power_heuristic(ppdf, lpdf));
                        }
                    }
                }
                f3 wi = s3(0.0f);
                float pdf = 0.0f, pw = 0.0f;
                const float u0 = sampler_next(sc, smp);
                const float u1 = sampler_next(sc, smp);
                if (mixed_phase_sample(vw, gs, nv, u0, u1, -ray_d, pw, wi, pdf)) {
                    beta = beta * pw;
                    const float rr = gmin(gmax(beta.x, gmax(beta.y, beta.z)), 1.0f);
                    bool cont = true;
                    if (depth + 1 > rs.ray_depth) {
                        if (rs.enable_russian_roulette && sampler_next(sc, smp) < rr) beta = beta / gmax(rr, kMinColorEpsilon);
                        else cont = false;
                    }
                    if (cont && depth + 1 <= kMaxRayDepth) {
                        alive = true;
                        next_o = vpos;
                        next_d = normalize(wi);
                        never_scatter = false;
                        depth += 1;
                    }
                }
            } else {
                const uint32_t inst = ha.y, tri = ha.z;
                const int back = (ha.w & kHitBack) ? 1 : 0;
                const DevInstance &in = sc.instances[inst];
                const uint32_t iflags = in.flags;
                const bool is_sphere = (iflags & kInstSphere) != 0;
                const ShadeTri &st = sc.shade[in.shade_base + (is_sphere ? 0u : tri)];
                const f3 b{hb.x, hb.y, hb.z};
                f3 p0 = s3(0.0f), p1 = s3(0.0f), p2 = s3(0.0f);
                f3 pos, perr, ng;
                const float radius = gmax(in.radius, kFloatEpsilon);
                if (is_sphere) {
                    // FetchSphereShadingData (src/ShadingUtilities.cuh:636-711)
                    pos = b;
                    perr = error_gamma(5) * abs3(b);
                    ng = b / radius;
                } else {
                    p0 = f3{st.p[0], st.p[1], st.p[2]};
                    p1 = f3{st.p[3], st.p[4], st.p[5]};
                    p2 = f3{st.p[6], st.p[7], st.p[8]};
                    const float xs = fabsf(p0.x * b.x) + fabsf(p1.x * b.y) + fabsf(p2.x * b.z);
                    const float ys = fabsf(p0.y * b.x) + fabsf(p1.y * b.y) + fabsf(p2.y * b.z);
                    const float zs = fabsf(p0.z * b.x) + fabsf(p1.z * b.y) + fabsf(p2.z * b.z);
                    pos = interp3(p0, p1, p2, b);
                    perr = error_gamma(7) * f3{xs, ys, zs};
                    ng = cross(p1 - p0, p2 - p0);
                }
                const DevMaterial &m = sc.materials[in.material];
                const uint32_t mtype = m.type;
                uint32_t prio;
                fetch_ior(m, &prio);
                // the medium records give the pass-through test and the two
                // IORs; nothing later needs the records themselves
                float ray_ior, ex_ior;
                bool pass_through;
                {
                    const Transfer t = load_transfer(ps, p, cap);
                    pass_through = (t.count != 0 && prio < t.get(t.max_idx).priority) || (iflags & kInstBoundary);
                    ray_ior = t.ray_ior();
                    ex_ior = t.ex_ior(inst);
                }
                if (pass_through || mtype == YMRT_MATERIAL_LIGHT) {
                    pos = xf_position(in.o2w, pos, &perr);
                    ng = normalize(xf_transpose_vector(in.w2o, ng));
                }
                if (pass_through) {
                    // pass through without consuming depth; the medium records
                    // change at the start of the next extend
                    pending = true;
                    alive = true;
                    next_o = offset_ray_origin(pos, perr, ray_d, ng);
                    next_d = normalize(ray_d);
                    next_li = inst;
                    next_lt = tri;
                } else if (mtype == YMRT_MATERIAL_LIGHT) {
                    if (never_scatter)
                        L = L + light_clamp(rs, beta * f3{m.light_color[0], m.light_color[1], m.light_color[2]} *
                                                    m.intensity,
                                            depth);
                } else {
                    f3 sn, dpdu, dpdv;
                    f2 uv;
                    if (is_sphere) {
                        sn = ng;
                        const float theta = ym_acosf(ng.y);
                        float phi = ym_atan2f(ng.z, ng.x);
                        if (phi < 0.0f) phi += kTwoPi;
                        uv = f2{0.5f * phi * kInvPi, theta * kInvPi};
                        float sin_theta = 0.0f;
                        if (1.0f - ng.y * ng.y > 0.0f) sin_theta = sqrtf(1.0f - ng.y * ng.y);
                        float cos_phi = 1.0f, sin_phi = 0.0f;
                        if (sin_theta > 0.0f) {
                            cos_phi = ng.x / sin_theta;
                            sin_phi = ng.z / sin_theta;
                        }
                        dpdu = f3{-b.z, 0.0f, b.x} * kTwoPi;
                        dpdv = f3{b.y * cos_phi, -radius * sin_theta, sin_phi * b.y} * kPi;
                    } else {
                        sn = (iflags & kInstNoNormals)
                                 ? ng
                                 : interp3(f3{st.n[0], st.n[1], st.n[2]}, f3{st.n[3], st.n[4], st.n[5]},
                                           f3{st.n[6], st.n[7], st.n[8]}, b);
                        const f2 uv0{st.uv[0], st.uv[1]}, uv1{st.uv[2], st.uv[3]}, uv2{st.uv[4], st.uv[5]};
                        uv = f2{uv0.x * b.x + uv1.x * b.y + uv2.x * b.z, uv0.y * b.x + uv1.y * b.y + uv2.y * b.z};
                        float du01 = uv1.x - uv0.x, dv01 = uv1.y - uv0.y, du02 = uv2.x - uv0.x, dv02 = uv2.y - uv0.y;
                        float det = du01 * dv02 - dv01 * du02;
                        if (fabsf(det) < 1E-8f) {
                            du01 = 1.0f;
                            dv01 = 0.0f;
                            du02 = 0.0f;
                            dv02 = 1.0f;
                            det = 1.0f;
                        }
                        const float idet = 1.0f / det;
                        const f3 dp01 = p1 - p0, dp02 = p2 - p0;
                        dpdu = (dv02 * dp01 - dv01 * dp02) * idet;
                        dpdv = (-du02 * dp01 + du01 * dp02) * idet;
                    }
                    const f3 pos_obj = pos;
                    pos = xf_position(in.o2w, pos, &perr);
                    sn = normalize(xf_transpose_vector(in.w2o, sn));
                    ng = normalize(xf_transpose_vector(in.w2o, ng));
                    dpdu = xf_vector(in.o2w, dpdu);
                    dpdv = xf_vector(in.o2w, dpdv);
                    const bool is_default = mtype == YMRT_MATERIAL_DEFAULT;
                    const uint32_t normal_tex = kGraph && is_default ? m.normal_mapping_tex : kEmpty;
                    const uint32_t bump_tex = kGraph && is_default && normal_tex == kEmpty ? m.bump_mapping_tex : kEmpty;
                    TexEval<kGraph> tx{sc, TexCoord{uv, pos, pos_obj}, tex_stk, (int)lane_id()};
                    float4 duv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (bump_tex != kEmpty)
                        duv = camera_uv_differentials(sc.camera, pos, sn, ng, dpdu, dpdv, rs.width, rs.height, 0.25f,
                                                      0.25f);
                    Bsdf bs;
                    init_shading_space(bs, tx, normal_tex, bump_tex, back, ray_d, ng, sn, dpdu, dpdv, in.w2o, duv);
                    bs.mask = 0;
                    if (is_default) init_default_bsdf<kClass>(tx, bs, m, ray_d, back, ray_ior, ex_ior);
                    // adjoint replay: an albedo or texel factor of this vertex
                    // scales every later deposit of the path (next-event ones
                    // included); a target roughness scales each next-event ray
                    // by its own BSDF value and the continuation by its own
                    int32_t rough_entry = -1;
                    float rough_r = 0.0f, adj_gc = 0.0f;
                    f3 adj_dl = s3(0.0f);
                    if constexpr (kAdj) {
                        const AdjParams &A = prm.adj;
                        const int32_t as = A.alb_slot[in.material];
                        const bool tex_on = A.tex != kEmpty && m.diffuse_albedo_tex == A.tex;
                        rough_entry = A.rough_slot[in.material];
                        if (as >= 0 || tex_on || rough_entry >= 0)
                            adj_dl = adj_pixel(prm, p) * (xyz(ps.out[p]) - L);
                        if (as >= 0) {
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const float al = m.diffuse_albedo[c];
                                if (al > 0.0f) sums.albedo((uint32_t)as, c, comp(adj_dl, c) / al);
                            }
                        }
                        if (tex_on) adj_texels(sc, A, uv, adj_dl);
                        if (rough_entry >= 0) rough_r = m.roughness_x;
                    }
                    const f3 wo = bs.to_local(-ray_d);
                    if (rs.enable_distant_light && sc.distant_count > 0) {
                        const DevDistant &dl = sc.distant[pick_distant(sc.distant_count, sampler_next(sc, smp))];
                        {
                            f3 ldir = s3(0.0f), Li = s3(0.0f);
                            float lpdf = 0.0f;
                            const float u0 = sampler_next(sc, smp);
                            const float u1 = sampler_next(sc, smp);
                            distant_sample_li(dl, u0, u1, ldir, Li, lpdf);
                            const f3 lwi = bs.to_local(ldir);
                            float wpdf = 0.0f;
                            const f3 f = bsdf_eval<kLobes>(bs, wo, lwi, wpdf);
                            if (maxcomp(f) > kMinColorEpsilon) {
                                const bool transmit = dot(-ray_d, ng) * dot(ldir, ng) < 0.0f;
                                const bool invalid = (!same_hemi(wo, lwi) && !transmit) || (same_hemi(wo, lwi) && transmit);
                                if (!invalid)
                                    so.emit(2, offset_ray_origin(pos, perr, ldir, ng), ldir, kTMax, transmit,
                                            min3(f, s3(kMaxColorClamp)), min3(Li, s3(kMaxColorClamp)),
                                            power_heuristic(lpdf, wpdf),
                                            rough_entry >= 0 ? ggx_rough_dlog(rough_r, wo, lwi) : 0.0f);
                            }
                        }
                        {
                            f3 bw = s3(0.0f), bwi = s3(0.0f);
                            float bpdf = 0.0f;
                            const float u0 = sampler_next(sc, smp);
                            const float u1 = sampler_next(sc, smp);
                            const float u2 = sampler_next(sc, smp);
                            if (bsdf_sample<kLobes>(bs, wo, bw, bwi, bpdf, u0, u1, u2)) {
                                const f3 ldir = bs.to_world(bwi);
                                float lpdf = 0.0f;
                                const f3 Li = distant_eval_li(dl, ldir, lpdf);
                                if (maxcomp(Li) > kMinColorEpsilon) {
                                    const bool transmit = dot(-ray_d, ng) * dot(ldir, ng) < 0.0f;
                                    const bool invalid =
                                        (!same_hemi(wo, bwi) && !transmit) || (same_hemi(wo, bwi) && transmit);
                                    if (!invalid)
                                        so.emit(3, offset_ray_origin(pos, perr, ldir, ng), ldir, kTMax, transmit,
                                                min3(bw, s3(kMaxColorClamp)), min3(Li, s3(kMaxColorClamp)),
                                                power_heuristic(bpdf, lpdf),
                                                rough_entry >= 0 ? ggx_rough_dlog(rough_r, wo, bwi) : 0.0f);
                                }
                            }
                        }
                    }
                    if (rs.enable_shape_light && sc.light_count > 0) {
                        const int li = sample_light_power(sc, sampler_next(sc, smp), sel_pdf);
                        const DevLightRec &light = sc.light_recs[li];
                        {
                            f3 ldir = s3(0.0f), lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f), Li = s3(0.0f);
                            float lpdf = 0.0f;
                            const float u0 = sampler_next(sc, smp);
                            const float u1 = sampler_next(sc, smp);
                            const bool valid = light_sample_li(light, pos, u0, u1, ldir, lpos, lerr, lng, Li, lpdf);
                            const f3 lwi = bs.to_local(ldir);
                            float wpdf = 0.0f;
                            const f3 f = bsdf_eval<kLobes>(bs, wo, lwi, wpdf);
                            if (maxcomp(f) > kMinColorEpsilon && valid) {
                                const bool transmit = dot(-ray_d, ng) * dot(ldir, ng) < 0.0f;
                                const bool invalid = (!same_hemi(wo, lwi) && !transmit) || (same_hemi(wo, lwi) && transmit);
                                if (!invalid)
                                    so.emit(0, offset_ray_origin(pos, perr, ldir, ng), ldir,
                                            shadow_tmax(pos, ldir, lpos, lerr, lng), transmit,
                                            min3(f, s3(kMaxColorClamp)), min3(Li, s3(kMaxColorClamp)),
                                            power_heuristic(lpdf, wpdf),
                                            rough_entry >= 0 ? ggx_rough_dlog(rough_r, wo, lwi) : 0.0f);
                            }
                        }
                        {
                            f3 bw = s3(0.0f), bwi = s3(0.0f);
                            float bpdf = 0.0f;
                            const float u0 = sampler_next(sc, smp);
                            const float u1 = sampler_next(sc, smp);
                            const float u2 = sampler_next(sc, smp);
                            if (bsdf_sample<kLobes>(bs, wo, bw, bwi, bpdf, u0, u1, u2)) {
                                const f3 ldir = bs.to_world(bwi);
                                f3 lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f);
                                float lpdf = 0.0f;
                                const f3 Li = light_eval_li(light, pos, ldir, lpos, lerr, lng, lpdf);
                                if (maxcomp(Li) > kMinColorEpsilon) {
                                    const bool transmit = dot(-ray_d, ng) * dot(ldir, ng) < 0.0f;
                                    const bool invalid =
                                        (!same_hemi(wo, bwi) && !transmit) || (same_hemi(wo, bwi) && transmit);
                                    if (!invalid)
                                        so.emit(1, offset_ray_origin(pos, perr, ldir, ng), ldir,
                                                shadow_tmax(pos, ldir, lpos, lerr, lng), transmit,
                                                min3(bw, s3(kMaxColorClamp)), min3(Li, s3(kMaxColorClamp)),
                                                power_heuristic(bpdf, lpdf),
                                                rough_entry >= 0 ? ggx_rough_dlog(rough_r, wo, bwi) : 0.0f);
                                }
                            }
                        }
                    }
                    f3 wi = s3(0.0f), bw = s3(0.0f);
                    float pdf = 0.0f;
                    const float u0 = sampler_next(sc, smp);
                    const float u1 = sampler_next(sc, smp);
                    const float u2 = sampler_next(sc, smp);
                    if (bsdf_sample<kLobes>(bs, wo, bw, wi, pdf, u0, u1, u2)) {
                        if constexpr (kAdj) {
                            if (rough_entry >= 0) {
                                adj_gc = ggx_rough_dlog(rough_r, wo, wi);
                                sums.roughness((uint32_t)rough_entry, (adj_dl.x + adj_dl.y + adj_dl.z) * adj_gc);
                            }
                        }
                        beta = beta * bw;
                        const float rr = gmin(gmax(beta.x, gmax(beta.y, beta.z)), 1.0f);
                        bool cont = true;
                        if (depth + 1 > rs.ray_depth) {
                            if (rs.enable_russian_roulette && sampler_next(sc, smp) < rr) beta = beta / gmax(rr, kMinColorEpsilon);
                            else cont = false;
                        }
                        if (cont) {
                            const f3 nd = bs.to_world(wi);
                            const bool transmit = dot(-ray_d, ng) * dot(nd, ng) < 0.0f;
                            const bool invalid = (!same_hemi(wo, wi) && !transmit) || (same_hemi(wo, wi) && transmit);
                            if (!invalid && depth + 1 <= kMaxRayDepth) {
                                pending = transmit;
                                alive = true;
                                next_o = offset_ray_origin(pos, perr, nd, ng);
                                next_d = normalize(nd);
                                next_li = inst;
                                next_lt = tri;
                                never_scatter = false;
                                depth += 1;
                            }
                        }
                    }
                    // the shadow stage subtracts the continuation's factor from
                    // each next-event ray's own
                    if constexpr (kAdj)
                        ps.adj_g[p] = make_uint2(rough_entry >= 0 ? (uint32_t)rough_entry : kEmpty,
                                                 __float_as_uint(adj_gc));
                }
            }
            // ---------------------------------------------- write back
            const bool nee = (so.flags & kShActiveMask) != 0;
            store_sampler(ps, p, smp);
            ps.L[p] = f4(L, 0.0f);
            if (nee) {
                out[0] = true;
                if (!alive) so.flags |= kShTerminate;
                ps.sh_flags[p] = so.flags;
                ps.sh_c[p] = f4(sh_beta, sel_pdf);
            } else if (!alive) {
                ps.out[p] = f4(L, 0.0f);
            }
            if (alive) {
                out[1] = true;
                ps.ray_o[p] = f4(next_o, 0.0f);
                ps.ray_d[p] = f4(next_d, 0.0f);
                ps.last[p] = make_uint2(next_li, next_lt);
                ps.beta[p] = f4(beta, 0.0f);
            }
            info = (uint32_t)depth | (never_scatter ? kInfoNeverScatter : 0u) | (pending ? kInfoPendingBunch : 0u) |
                   (vertex << 16);
            ps.info[p] = info;
        }
    });
    sums.end();
}


template <bool kVolume, bool kGraph, uint32_t kClass = kMatGeneral, bool kAdj = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kGraph ? 2 : 3))) void k_shade_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass)
{
    constexpr uint32_t kQueue = kVolume ? kCtrHitVol : kClass == kMatDiffuse ? kCtrHitDiffuse
                                : kClass == kMatMetal ? kCtrHitMetal : kCtrHit;
    constexpr uint32_t kWork = kVolume ? kCtrWorkShadeVol : kClass == kMatDiffuse ? kCtrWorkShadeDiffuse
                               : kClass == kMatMetal ? kCtrWorkShadeMetal : kCtrWorkShade;
    for_slots(
        fs, nfs, [&](const FrameSlot &f) { return pass_counter(f, pass, kQueue); },
        [&](const FrameSlot &f) { return pass_drained(f, pass, kWork, kQueue); },
        [&](const FrameSlot &f) { shade_body<kVolume, kGraph, kClass, kAdj>(f.sc, f.prm, f.ps, pass); });
}



// ------------------------------------------------------------ shadow
// NEE contribution of a path's shadow rays: per light kind, the sum of
// mis * tr * f * Li over the rays that reached the light, divided by the
// light-selection pdf, clamped and deposited, distant light first
// (EvalDistantLight, src/Shading.cu:83-198, then EvalShapeLight 294-417, as
// the shading stage adds them at 2217-2240).  with_tr: the transmittance of
// each ray is in sh_tr (paths in media), otherwise 1.  kAdj (the adjoint's
// replay): at a surface vertex with a target roughness, each ray's deposit D
// adds adj . D * (g_ray - g_continuation) to that roughness's gradient (the
// continuation's event counted D with the continuation's factor).
template <bool kAdj = false>
__device__ __forceinline__ void shadow_finish(const PtParams &prm, const PathState &ps, uint32_t p, uint32_t fl,
                                              uint32_t distant_count, bool with_tr, AdjSums<kAdj> *sums = nullptr)
{
    const uint32_t cap = prm.cap;
    f3 shape = s3(0.0f), dist = s3(0.0f);
    f3 rc[kShadowSlots];
    float rg[kShadowSlots];
    static_for<0, kShadowSlots>([&](auto K) {
        rc[K()] = s3(0.0f);
        rg[K()] = 0.0f;
        if ((fl & (1u << K())) && (fl & (kShClear0 << K()))) {
            const size_t o = (size_t)K() * cap + p;
            const float4 f4v = ps.sh_f[o];
            const float4 li4 = ps.sh_li[o];
            const f3 tr = with_tr ? xyz(ps.sh_tr[o]) : s3(1.0f);
            const f3 c = ((f4v.w * tr) * xyz(f4v)) * xyz(li4);
            if (K() < 2) shape = shape + c;
            else dist = dist + c;
            if constexpr (kAdj) {
                rc[K()] = c;
                rg[K()] = li4.w;
            }
        }
    });
    const float4 c4 = ps.sh_c[p];
    const f3 beta = xyz(c4);
    if constexpr (kAdj) {
        if (!(fl & kShVolume)) {
            const uint2 ag = ps.adj_g[p];
            if (ag.x != kEmpty) {
                const float gc = __uint_as_float(ag.y);
                const f3 a = adj_pixel(prm, p);
                float acc = 0.0f;
                static_for<0, kShadowSlots>([&](auto K) {
                    const float scale = K() < 2 ? safe_rcp(c4.w) : (float)distant_count;
                    const f3 d = beta * (rc[K()] * scale);
                    acc += dot(a, d) * (rg[K()] - gc);
                });
                sums->roughness(ag.x, acc);
            }
        }
    }
    const int depth = (int)((fl >> kShDepthShift) & 0xFFu);
    f3 L = xyz(ps.L[p]);
    if (distant_count > 0) L = L + light_clamp(prm.rs, beta * (dist / (1.0f / (float)distant_count)), depth);
    // EvalShapeLight returns direct * SafeRcp(select pdf); the volume variant
    // divides when the pdf is positive.
    const f3 sel = (fl & kShVolume) ? (c4.w > 0.0f ? shape / c4.w : s3(0.0f)) : shape * safe_rcp(c4.w);
    L = L + light_clamp(prm.rs, beta * sel, depth);
    ps.L[p] = f4(L, 0.0f);
    if (fl & kShTerminate) ps.out[p] = f4(L, 0.0f);
}

// Any-hit for each shadow ray against non-boundary instances
// (BVHTraverseShadow, src/ShadingUtilities.cuh:563-635).  In scenes with media,
// paths with an unoccluded ray go on to k_transmit for the transmittance along
// it; the others are finished here.
template <bool kMedia, bool kSphere, int kStack, bool kAdj = false>
__device__ __forceinline__ void shadow_body(const DevScene &sc, const PtParams &prm, const PathState &ps, uint32_t pass)
{
    __shared__ uint32_t stack[kStackWindow * kBlock];
    __shared__ float pre_park[kPreSlots * kBlock];
    __shared__ float adj_lds[kAdj ? kAdjLds : 1];
    AdjSums<kAdj> sums{adj_lds, &prm.adj};
    sums.begin();
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    const uint32_t cap = prm.cap;
    LaneScratch<kBlock, kStackWindow, true> ls{stack, nullptr, nullptr, nullptr, (int)lane_id(), ps.spill,
                                               ps.spill_threads, blockIdx.x * kBlock + threadIdx.x, pre_park};
    uint32_t nsh = 0;
    TraceCounters tc{0, 0};
    RefillRange<16> rr{cnt + kCtrWorkShadow, cnt[kCtrShadow], 0u, 0u, true};

    // one lane walks its path's recorded rays (slots k) in order
    Traversal<kModeSkipBoundary, true, false, true, kSphere, kStack> tv;
    uint32_t p = 0, fl = 0;
    int k = 0;
    int r = kLaneIdle;
    auto start_ray = [&]() {
        const float4 o4 = gload(ps.sh_o + (size_t)k * cap + p);
        const float4 d4 = gload(ps.sh_d + (size_t)k * cap + p);
        Query q;
        q.o = xyz(o4);
        q.d = xyz(d4);
        q.tmin = kTMin;
        q.tmax = o4.w;
        q.last_inst = kEmpty;
        q.last_tri = kEmpty;
        r = tv.init(sc, q);
    };
    for (;;) {
        uint32_t item = 0;
        if (rr.more && rr.refill(r == kLaneIdle, item)) {
            p = gload(ps.q_shadow + item);
            fl = gload(ps.sh_flags + p);
            k = __builtin_ctz(fl & kShActiveMask);  // a queued path records at least one ray
            start_ray();
        }
        const uint32_t want = rr.want();
        for (;;) {
            const uint64_t busy = __ballot(r == kTravContinue);
            if (busy == 0 || (uint32_t)(kBlock - __popcll(busy)) >= want) break;
            wave_step<24, false>(tv, r, sc, ls, tc);
        }
        // finished rays: note the result, start the path's next ray or
        // finish the path (a ray can end in init, hence the loop)
        while (r != kTravContinue && r != kLaneIdle) {
            ++nsh;
            if (r == kTravOverflow) atomicOr(prm.error, 1u);
            if (r == kTravMiss) fl |= kShClear0 << k;
            const uint32_t rest = fl & kShActiveMask & ~((2u << k) - 1u);
            if (rest) {
                k = __builtin_ctz(rest);
                start_ray();
            } else {
                r = kLaneIdle;
                const bool walk = kMedia && prm.media && (fl & kShClearMask);
                if (walk) {
                    ps.sh_flags[p] = fl;
                    ps.flag[p] = (pass << kFlagBits) | kFlagTransmit;
                } else {
                    shadow_finish<kAdj>(prm, ps, p, fl, sc.distant_count, false, &sums);
                }
            }
        }
        if (!rr.more && __ballot(r != kLaneIdle) == 0) break;
    }
    sums.end();
    wave_add64(prm.stats + kStatShadow, nsh);
    wave_add64(prm.stats + kStatShadowNodes, tc.nodes);
    wave_add64(prm.stats + kStatShadowTris, tc.tris);
}


template <bool kMedia, bool kSphere, int kStack, bool kAdj = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_shadow_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass)
{
    for_slots(
        fs, nfs, [&](const FrameSlot &f) { return pass_counter(f, pass, kCtrShadow); },
        [&](const FrameSlot &f) { return pass_drained(f, pass, kCtrWorkShadow, kCtrShadow); },
        [&](const FrameSlot &f) { shadow_body<kMedia, kSphere, kStack, kAdj>(f.sc, f.prm, f.ps, pass); });
}



// Transmittance of the unoccluded shadow rays of paths in media: TraceTr
// (src/VolumeEval.cuh:420-468) walks the boundary instances along the ray,
// updating the medium records at each crossing and multiplying the segment
// transmittances (EvalTr, 253-418).
// kHet as k_medium's.  kAdj (the adjoint's replay): the shadow_finish
// roughness terms, and with a target grid volume each ray's ratio-tracking
// events, found by walking the ray a second time with the same random
// numbers once its deposit is known.
template <bool kSphere, int kStack, int kHet, bool kAdj = false>
__device__ __forceinline__ void transmit_body(const DevScene &sc, const PtParams &prm, const PathState &ps, uint32_t pass)
{
    __shared__ float tex_stk[kHet == 2 ? kTexStack * 3 * kBlock : 1];
    __shared__ uint32_t stack[kStackWindow * kBlock];
    __shared__ uint32_t nbp[kNearbyCap * kBlock];
    __shared__ float nbt[kNearbyCap * kBlock];
    __shared__ int nbb[kNearbyCap * kBlock];
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    const uint32_t n = cnt[kCtrTransmit];
    const uint32_t cap = prm.cap;
    LaneScratch<kBlock, kStackWindow> ls{stack, nbp, nbt, nbb, (int)lane_id(), ps.spill, ps.spill_threads,
                                         blockIdx.x * kBlock + threadIdx.x};
    uint32_t ntr = 0;
    TraceCounters tc{0, 0};
    __shared__ float adj_lds[kAdj ? kAdjLds : 1];
    AdjSums<kAdj> sums{adj_lds, &prm.adj};
    sums.begin();
    run_items<0>(cnt + kCtrWorkTransmit, n, nullptr, nullptr, [&](uint32_t i, uint32_t &slot, bool(&)[1]) {
        {
            const uint32_t p = ps.q_transmit[i];
            slot = p;
            const uint32_t fl = ps.sh_flags[p];
            const uint32_t info = ps.info[p];
            const uint4 ha = ps.hit_a[p];
            HashRng rng;
            {
                // 64-bit like k_raygen: a frame may hold more than 2^32 path slots
                const uint64_t gp = prm.chunk_base + p;
                const uint32_t spp = (uint32_t)prm.rs.spp;
                const uint64_t vpix = gp / spp;
                const uint32_t s = (uint32_t)(gp - vpix * spp);
                const uint32_t tile = (uint32_t)(vpix >> 8), within = (uint32_t)(vpix & 255u);
                const uint32_t ty = tile / prm.tiles_x, tx = tile - ty * prm.tiles_x;
                const uint32_t px = tx * 16u + (within & 15u), py = ty * 16u + (within >> 4);
                rng.key[0] = py * prm.rs.width + px;
                rng.key[1] = info_vertex(info);
                rng.key[2] = s;
                rng.key[3] = (uint32_t)prm.frame;
                rng.ctr = 0;
            }
            // distant-light rays first, as the shading stage evaluates them
#pragma unroll 1
            for (int j = 0; j < kShadowSlots; ++j) {
                const int k = (j + 2) & 3;
                if (!(fl & (1u << k)) || !(fl & (kShClear0 << k))) continue;
                auto walk = [&](auto &&ev) -> f3 {
                    const float4 o4 = ps.sh_o[(size_t)k * cap + p];
                    const float4 d4 = ps.sh_d[(size_t)k * cap + p];
                    f3 tr = s3(1.0f);
                    // the vertex's records and nearby list are reloaded per ray
                    // rather than held across the loop (fewer live registers)
                    Transfer t = load_transfer(ps, p, cap);
                    if (fl & (kShTransmit0 << k)) boundary_bunch(sc, t, load_nearby(ps, p, cap, ha.w));
                    const bool volv = (fl & kShVolume) != 0;
                    Query tq;
                    tq.o = xyz(o4);
                    tq.d = xyz(d4);
                    tq.tmin = kTMin;
                    tq.tmax = o4.w;
                    tq.last_inst = volv ? kEmpty : ha.y;
                    tq.last_tri = volv ? kEmpty : ha.z;
                    const float maxd = o4.w;
                    float traced = kTMin;
                    int guard = 0;
                    while (traced < maxd) {
                        Hit bh;
                        bh.back = 0;
                        int bnc = 0;
                        int br = 0;
                        ++ntr;
                        if (sc.has_boundaries)
                            br = traverse<kModeBoundaryOnly, false, true, true, kSphere, kStack>(sc, tq, bh, bnc, ls, tc);
                        if (br < 0) {
                            atomicOr(prm.error, 1u);
                            break;
                        }
                        if (br == 0) {
                            tr = tr * eval_tr<kHet>(sc, tq.o, tq.d, t, gmax(maxd - traced, 0.0f), rng, tex_stk, ev);
                            break;
                        }
                        tr = tr * eval_tr<kHet>(sc, tq.o, tq.d, t, bh.t, rng, tex_stk, ev);
                        traced += bh.t;
                        f3 bp, bperr, bng;
                        hit_geometry(sc, bh.inst, bh.tri, bh.b0, bh.b1, bh.b2, bp, bperr, bng);
                        NearbyList bnb;
                        bnb.n = bnc;
                        static_for<0, kNearbyMax>([&](auto J) {
                            bnb.prim[J()] = nbp[J() * kBlock + lane_id()];
                            bnb.t[J()] = nbt[J() * kBlock + lane_id()];
                            bnb.back[J()] = nbb[J() * kBlock + lane_id()];
                        });
                        boundary_bunch(sc, t, bnb);
                        // the next segment is a new Ray, direction renormalised
                        tq.o = offset_ray_origin(bp, bperr, tq.d, bng);
                        tq.d = normalize(tq.d);
                        tq.tmax = kTMax;
                        tq.last_inst = bh.inst;
                        tq.last_tri = bh.tri;
                        if (++guard > 64) break;
                    }
                    return tr;
                };
                const uint32_t ctr0 = rng.ctr;
                const f3 tr = walk(NoTrackEvents{});
                ps.sh_tr[(size_t)k * cap + p] = f4(tr, 0.0f);
                if constexpr (kAdj && kHet > 0) {
                    if (prm.adj.vol != kEmpty) {
                        // this ray's deposit, as shadow_finish sums it
                        const size_t o = (size_t)k * cap + p;
                        const float4 f4v = ps.sh_f[o];
                        const float4 li4 = ps.sh_li[o];
                        const float4 c4 = ps.sh_c[p];
                        const f3 c = ((f4v.w * tr) * xyz(f4v)) * xyz(li4);
                        const float scale = k >= 2 ? (float)sc.distant_count
                                            : (fl & kShVolume) ? (c4.w > 0.0f ? 1.0f / c4.w : 0.0f)
                                                               : safe_rcp(c4.w);
                        const float d = dot(adj_pixel(prm, p), xyz(c4) * (c * scale));
                        if (d != 0.0f) {
                            const uint32_t ctr1 = rng.ctr;
                            rng.ctr = ctr0;
                            AdjTrackEvents ev = adj_track(sc, prm.adj, d, kAdjRayEvents);
                            (void)walk(ev);
                            rng.ctr = ctr1;
                        }
                    }
                }
            }
            shadow_finish<kAdj>(prm, ps, p, fl, sc.distant_count, true, &sums);
        }
    });
    sums.end();
    wave_add64(prm.stats + kStatTransmittance, ntr);
    wave_add64(prm.stats + kStatTransmitNodes, tc.nodes);
    wave_add64(prm.stats + kStatTransmitTris, tc.tris);
}


template <bool kSphere, int kStack, int kHet, bool kAdj = false>
__global__ __launch_bounds__(kBlock) void k_transmit_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass)
{
    for_slots(
        fs, nfs, [&](const FrameSlot &f) { return pass_counter(f, pass, kCtrTransmit); },
        [&](const FrameSlot &f) { return pass_drained(f, pass, kCtrWorkTransmit, kCtrTransmit); },
        [&](const FrameSlot &f) { transmit_body<kSphere, kStack, kHet, kAdj>(f.sc, f.prm, f.ps, pass); });
}



// ------------------------------------------------------------ flush
// Paths still queued after the last pass end with the radiance they carry.
__device__ __forceinline__ void flush_body(const PtParams &prm, const PathState &ps, uint32_t pass)
{
    uint32_t *cnt = prm.counters + pass * kCtrStride;
    const uint32_t n = cnt[kCtrActive];
    const uint32_t *q = ps.queue[pass & 1];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t p = q[i];
        ps.out[p] = ps.L[p];
    }
}


// one row of the grid (blockIdx.y) per slot
__global__ __launch_bounds__(kBlock) void k_flush_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass)
{
    const FrameSlot &f = fs[blockIdx.y];
    flush_body(f.prm, f.ps, pass);
}



// ------------------------------------------------------------ resolve
__device__ __forceinline__ void resolve_body(const PtParams &prm, const PathState &ps, float4 *noise)
{
    const uint32_t spp = (uint32_t)prm.rs.spp;
    const uint32_t pix_in_chunk = prm.chunk_paths / spp;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= pix_in_chunk) return;
    const uint64_t vpix = prm.chunk_base / spp + k;
    const uint32_t tile = (uint32_t)(vpix >> 8), within = (uint32_t)(vpix & 255u);
    const uint32_t ty = tile / prm.tiles_x, tx = tile - ty * prm.tiles_x;
    const uint32_t px = tx * 16u + (within & 15u), py = ty * 16u + (within >> 4);
    if (px >= prm.rs.width || py >= prm.rs.height) return;
    const float inv = 1.0f / (float)(prm.rs.spp > 1 ? prm.rs.spp : 1);
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    const float4 *o = ps.out + (size_t)k * spp;
    for (uint32_t s = 0; s < spp; ++s) {
        const float4 v = o[s];
        a0 += v.x * inv;
        a1 += v.y * inv;
        a2 += v.z * inv;
        a3 += inv;
    }
    noise[(size_t)py * prm.rs.width + px] = make_float4(a0, a1, a2, a3);
    // the id image holds sample 0's camera-ray instance (RayTraceEditorView,
    // src/Shading.cu:1545-1547, keeps whichever sample wrote last)
    if (prm.aov) prm.aov[(size_t)py * prm.rs.width + px] = ps.first_hit[(size_t)k * spp];
}


// one row of the grid (blockIdx.y) per slot, each into its frame's image
__global__ void k_resolve_m(const FrameSlot *__restrict__ fs)
{
    const FrameSlot &f = fs[blockIdx.y];
    resolve_body(f.prm, f.ps, f.prm.noise);
}



// ------------------------------------------------------------ post
__global__ void k_accumulate(const float4 *noise, float4 *accum, uint32_t n, int frame)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = noise[i];
    if (frame == 1) {
        accum[i] = v;
        return;
    }
    const float4 a = accum[i];
    const float fm1 = (float)(frame - 1), fr = (float)frame;
    accum[i] = make_float4((a.x * fm1 + v.x) / fr, (a.y * fm1 + v.y) / fr, (a.z * fm1 + v.z) / fr,
                           (a.w * fm1 + v.w) / fr);
}

__global__ void k_tonemap(float4 *img, uint32_t n, float gamma, float exposure)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = img[i];
    float c[3] = {v.x, v.y, v.z};
    for (int k = 0; k < 3; ++k) {
        const float x = c[k] * exposure;
        const float y = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
        c[k] = ym_powf(y, 1.0f / gamma);
    }
    img[i] = make_float4(c[0], c[1], c[2], 1.0f);
}

// ------------------------------------------------------------ launchers
// Persistent pass kernels: the caller's persistent_waves, or by default one
// wave per 256 paths of the launch (all slots), within [1024, kMaxPersistentWaves] (the
// traversal spill area holds kSpillThreads threads): large chunks fill 256
// CUs several times over, and small frames launch fewer waves, which drain
// sooner (256x256 at 4 spp: 7.8 ms with 1024 waves against 9.9 ms with one
// path per lane); never more waves than the items need.  Every pass kernel
// fetches work dynamically or strides over the grid, so the count changes
// timing only.
uint32_t persistent_waves(uint64_t paths, uint32_t requested, uint32_t spill_threads)
{
    static_assert(kMaxPersistentWaves * kBlock == kSpillThreads, "one spill row per lane");
    const uint64_t waves = (paths + kBlock - 1) / kBlock;
    const uint32_t most = spill_threads / kBlock;  // one spill row per lane
    uint64_t cap = (paths + 255u) / 256u;
    cap = cap < 1024u ? 1024u : cap;
    if (requested) cap = requested;
    cap = cap < most ? cap : most;
    return (uint32_t)(waves < cap ? (waves ? waves : 1) : cap);
}

// ------------------------------------------------------------ multi-frame launchers
// One launch per stage for every slot of a round (ymrt_render_batch).  The
// variant is the union of the slots' scene features (each feature's code is
// inert where a scene lacks it); waves: the persistent grid shared by all
// slots; blocks: the most compaction / resolve blocks any slot needs.
__global__ __launch_bounds__(kBlock) void k_active_m(const FrameSlot *__restrict__ fs, uint32_t nfs, uint32_t pass,
                                                     uint32_t *out)
{
    uint32_t v = 0;
    for (uint32_t s = threadIdx.x; s < nfs; s += kBlock) v += fs[s].prm.counters[pass * kCtrStride + kCtrActive];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (threadIdx.x == 0) *out = v;
}

// A persistent pass kernel gets at most as many one-wave blocks as the GPU
// holds at once (its occupancy times the CU count, queried once per
// kernel): further waves could only start once the stage's work has run
// out, and each would still take one fetch atomic on its counter.
template <typename... Params, typename... Args>
static void launch_persistent(void (*kernel)(Params...), uint32_t waves, hipStream_t s, Args... args)
{
    static std::mutex mu;
    static std::unordered_map<const void *, uint32_t> resident;
    uint32_t cap = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = resident.find((const void *)kernel);
        if (it == resident.end()) {
            int blocks = 0, dev = 0, cus = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kBlock, 0) == hipSuccess &&
                hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
                cap = (uint32_t)(blocks * cus);
            resident.emplace((const void *)kernel, cap);
        } else {
            cap = it->second;
        }
    }
    const uint32_t grid = cap > 0 && cap < waves ? cap : waves;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, s, args...);
}

int launch_multi_active(const FrameSlot *fs, uint32_t nfs, uint32_t pass, uint32_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_active_m, dim3(1), dim3(kBlock), 0, s, fs, nfs, pass, out);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_raygen(const FrameSlot *fs, uint32_t nfs, uint32_t waves, hipStream_t s)
{
    launch_persistent(k_raygen_m, waves, s, fs, nfs);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

template <bool kNearby, bool kMedia, bool kSphere>
static void extend_m(const FrameSlot *fs, uint32_t nfs, uint32_t waves, uint32_t pass, hipStream_t s)
{
    launch_persistent((k_extend_m<kNearby, kMedia, kSphere, kMaxStackDepth>), waves, s, fs, nfs, pass);
}

int launch_multi_extend(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                        hipStream_t s)
{
    const int k = (v.nearby ? 4 : 0) | (v.media ? 2 : 0) | (v.sphere ? 1 : 0);
    switch (k) {
    case 0: extend_m<false, false, false>(fs, nfs, waves, pass, s); break;
    case 1: extend_m<false, false, true>(fs, nfs, waves, pass, s); break;
    case 2: extend_m<false, true, false>(fs, nfs, waves, pass, s); break;
    case 3: extend_m<false, true, true>(fs, nfs, waves, pass, s); break;
    case 4: extend_m<true, false, false>(fs, nfs, waves, pass, s); break;
    case 5: extend_m<true, false, true>(fs, nfs, waves, pass, s); break;
    case 6: extend_m<true, true, false>(fs, nfs, waves, pass, s); break;
    default: extend_m<true, true, true>(fs, nfs, waves, pass, s); break;
    }
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_medium(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                        hipStream_t s)
{
    if (v.adj && v.het == 2) launch_persistent((k_medium_m<2, true>), waves, s, fs, nfs, pass);
    else if (v.adj && v.het == 1) launch_persistent((k_medium_m<1, true>), waves, s, fs, nfs, pass);
    else if (v.het == 2) launch_persistent(k_medium_m<2>, waves, s, fs, nfs, pass);
    else if (v.het == 1) launch_persistent(k_medium_m<1>, waves, s, fs, nfs, pass);
    else launch_persistent(k_medium_m<0>, waves, s, fs, nfs, pass);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_compact(const FrameSlot *fs, uint32_t nfs, uint32_t blocks, uint32_t pass, int mode, hipStream_t s)
{
    const dim3 g(blocks ? blocks : 1, nfs), b(kCompactThreads), g1(4, nfs), b1(kScanThreads);
    if (mode == kCompactShade) {
        hipLaunchKernelGGL(k_compact_count_m<kCompactShade>, g, b, 0, s, fs, pass);
        hipLaunchKernelGGL(k_compact_scan_m<kCompactShade>, g1, b1, 0, s, fs);
        hipLaunchKernelGGL(k_compact_scatter_m<kCompactShade>, g, b, 0, s, fs, pass);
    } else if (mode == kCompactMedium) {
        hipLaunchKernelGGL(k_compact_count_m<kCompactMedium>, g, b, 0, s, fs, pass);
        hipLaunchKernelGGL(k_compact_scan_m<kCompactMedium>, g1, b1, 0, s, fs);
        hipLaunchKernelGGL(k_compact_scatter_m<kCompactMedium>, g, b, 0, s, fs, pass);
    } else {
        hipLaunchKernelGGL(k_compact_count_m<kCompactTransmit>, g, b, 0, s, fs, pass);
        hipLaunchKernelGGL(k_compact_scan_m<kCompactTransmit>, g1, b1, 0, s, fs);
        hipLaunchKernelGGL(k_compact_scatter_m<kCompactTransmit>, g, b, 0, s, fs, pass);
    }
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

template <bool kGraph, bool kAdj>
static void shade_m(const FrameSlot *fs, uint32_t nfs, uint32_t waves, uint32_t pass, hipStream_t s)
{
    launch_persistent((k_shade_m<false, kGraph, kMatGeneral, kAdj>), waves, s, fs, nfs, pass);
    launch_persistent((k_shade_m<false, kGraph, kMatDiffuse, kAdj>), waves, s, fs, nfs, pass);
    launch_persistent((k_shade_m<false, kGraph, kMatMetal, kAdj>), waves, s, fs, nfs, pass);
}

int launch_multi_shade(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, bool volume, uint32_t waves,
                       uint32_t pass, hipStream_t s)
{
    // volume vertices carry no adjoint events of their own
    if (volume) launch_persistent((k_shade_m<true, false>), waves, s, fs, nfs, pass);
    else if (v.graph) v.adj ? shade_m<true, true>(fs, nfs, waves, pass, s) : shade_m<true, false>(fs, nfs, waves, pass, s);
    else v.adj ? shade_m<false, true>(fs, nfs, waves, pass, s) : shade_m<false, false>(fs, nfs, waves, pass, s);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_shadow(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                        hipStream_t s)
{
    if (v.adj) {
        if (v.sphere) {
            if (v.media) launch_persistent((k_shadow_m<true, true, kMaxStackDepth, true>), waves, s, fs, nfs, pass);
            else launch_persistent((k_shadow_m<false, true, kMaxStackDepth, true>), waves, s, fs, nfs, pass);
        } else {
            if (v.media) launch_persistent((k_shadow_m<true, false, kMaxStackDepth, true>), waves, s, fs, nfs, pass);
            else launch_persistent((k_shadow_m<false, false, kMaxStackDepth, true>), waves, s, fs, nfs, pass);
        }
    } else if (v.sphere) {
        if (v.media) launch_persistent((k_shadow_m<true, true, kMaxStackDepth>), waves, s, fs, nfs, pass);
        else launch_persistent((k_shadow_m<false, true, kMaxStackDepth>), waves, s, fs, nfs, pass);
    } else {
        if (v.media) launch_persistent((k_shadow_m<true, false, kMaxStackDepth>), waves, s, fs, nfs, pass);
        else launch_persistent((k_shadow_m<false, false, kMaxStackDepth>), waves, s, fs, nfs, pass);
    }
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_transmit(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                          hipStream_t s)
{
    if (v.adj) {
        if (v.sphere) {
            if (v.het == 2) launch_persistent((k_transmit_m<true, kMaxStackDepth, 2, true>), waves, s, fs, nfs, pass);
            else if (v.het == 1) launch_persistent((k_transmit_m<true, kMaxStackDepth, 1, true>), waves, s, fs, nfs, pass);
            else launch_persistent((k_transmit_m<true, kMaxStackDepth, 0, true>), waves, s, fs, nfs, pass);
        } else {
            if (v.het == 2) launch_persistent((k_transmit_m<false, kMaxStackDepth, 2, true>), waves, s, fs, nfs, pass);
            else if (v.het == 1) launch_persistent((k_transmit_m<false, kMaxStackDepth, 1, true>), waves, s, fs, nfs, pass);
            else launch_persistent((k_transmit_m<false, kMaxStackDepth, 0, true>), waves, s, fs, nfs, pass);
        }
    } else if (v.sphere) {
        if (v.het == 2) launch_persistent((k_transmit_m<true, kMaxStackDepth, 2>), waves, s, fs, nfs, pass);
        else if (v.het == 1) launch_persistent((k_transmit_m<true, kMaxStackDepth, 1>), waves, s, fs, nfs, pass);
        else launch_persistent((k_transmit_m<true, kMaxStackDepth, 0>), waves, s, fs, nfs, pass);
    } else {
        if (v.het == 2) launch_persistent((k_transmit_m<false, kMaxStackDepth, 2>), waves, s, fs, nfs, pass);
        else if (v.het == 1) launch_persistent((k_transmit_m<false, kMaxStackDepth, 1>), waves, s, fs, nfs, pass);
        else launch_persistent((k_transmit_m<false, kMaxStackDepth, 0>), waves, s, fs, nfs, pass);
    }
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_flush(const FrameSlot *fs, uint32_t nfs, uint32_t pass, hipStream_t s)
{
    hipLaunchKernelGGL(k_flush_m, dim3(64, nfs), dim3(kBlock), 0, s, fs, nfs, pass);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int launch_multi_resolve(const FrameSlot *fs, uint32_t nfs, uint32_t pixel_blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_resolve_m, dim3(pixel_blocks ? pixel_blocks : 1, nfs), dim3(256), 0, s, fs);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}


// Texture graph at n given coordinates: coords[i*8 + 0..7] = (s, t,
// p_world.xyz, p_object.xyz), out[i*3 + 0..2] = the texture's rgb.
__global__ __launch_bounds__(kBlock) void k_texture_eval(DevScene sc, uint32_t tex, const float *coords, uint32_t n,
                                                         float *out)
{
    __shared__ float tex_stk[kTexStack * 3 * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float *c = coords + (size_t)i * 8;
    TexEval<true> tx{sc, TexCoord{f2{c[0], c[1]}, f3{c[2], c[3], c[4]}, f3{c[5], c[6], c[7]}}, tex_stk,
                     (int)threadIdx.x};
    const f3 v = tx.eval(tex);
    out[(size_t)i * 3 + 0] = v.x;
    out[(size_t)i * 3 + 1] = v.y;
    out[(size_t)i * 3 + 2] = v.z;
}

int launch_texture_eval(const DevScene &sc, uint32_t tex, const float *coords, uint32_t n, float *out, hipStream_t s)
{
    YM_REQUIRE(tex < sc.texture_count, "texture index out of range");
    if (n == 0) return YMRT_OK;
    hipLaunchKernelGGL(k_texture_eval, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, sc, tex, coords, n, out);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int accumulate(const float *noise, float *accum, uint32_t w, uint32_t h, int frame, hipStream_t s)
{
    const uint32_t n = w * h;
    if (n == 0) return YMRT_OK;
    hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, reinterpret_cast<const float4 *>(noise),
                       reinterpret_cast<float4 *>(accum), n, frame);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

int tonemap(float *rgba, uint32_t w, uint32_t h, float gamma, float exposure, hipStream_t s)
{
    const uint32_t n = w * h;
    if (n == 0) return YMRT_OK;
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, reinterpret_cast<float4 *>(rgba), n, gamma,
                       exposure);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

}  // namespace ym

// render.cpp -- host driver of one path-traced frame.
//
// Replaces RenderModule::PathTracingExecuteTask (src/core/RenderModule.cpp:494-653).
// The reference regenerates paths into a 1M-ray queue and round-trips three
// counters through the host after every stage (8 synchronisations per batch).
// Here the frame is cut into chunks of whole pixels (capacity paths each, 4M
// by default: ~1.2 GB of state, a sliver of the 288 GB HBM); each chunk runs
// raygen, then extend/shade/shadow passes whose queue lengths live in device
// counters, then a resolve into the noise image.  The host reads one counter
// per chunk, after ray_depth + 1 passes, to decide whether boundary
// pass-throughs need more passes.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "buffers.hpp"
#include "internal.hpp"
#include "pt.hpp"

namespace ym {

uint32_t persistent_waves(uint64_t paths, uint32_t requested, uint32_t spill_threads);
int launch_multi_raygen(const FrameSlot *fs, uint32_t nfs, uint32_t waves, hipStream_t s);
int launch_multi_extend(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                        hipStream_t s);
int launch_multi_medium(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                        hipStream_t s);
// mode: 0 free flight, 1 shading, 2 transmittance
int launch_multi_compact(const FrameSlot *fs, uint32_t nfs, uint32_t blocks, uint32_t pass, int mode, hipStream_t s);
int launch_multi_shade(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, bool volume, uint32_t waves,
                       uint32_t pass, hipStream_t s);
int launch_multi_shadow(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                        hipStream_t s);
int launch_multi_transmit(const FrameSlot *fs, uint32_t nfs, const MultiVariant &v, uint32_t waves, uint32_t pass,
                          hipStream_t s);
int launch_multi_flush(const FrameSlot *fs, uint32_t nfs, uint32_t pass, hipStream_t s);
int launch_multi_resolve(const FrameSlot *fs, uint32_t nfs, uint32_t pixel_blocks, hipStream_t s);

// Paths in flight per chunk when the settings leave it to the library.  Each
// chunk ends in passes that hold only a frame's longest paths, where the GPU
// idles; larger chunks amortise those tails (interior 1080p x 64 spp: 1417
// ms per frame in 4 M-path chunks, 1188 in 16 M, 1129 in 64 M).  The path
// state costs about kPathBytes per path, so a chunk is also held to a quarter
// of the device memory free when the buffers are first allocated, and never
// to more than the frame needs.
constexpr uint32_t kDefaultCapacity = 1u << 26;
constexpr uint64_t kPathBytes = 720;

void destroy_render_buffers(ymrt_scene *sc)
{
    delete sc->render;
    sc->render = nullptr;
}

void free_render_buffers(RenderBuffers *rb) { delete rb; }

// Capacity choices read the free device memory and then allocate; threads
// of a batch would otherwise all see the same free memory.
static std::mutex g_alloc_mu;

static int run_frame(ymrt_scene *sc, RenderBuffers *rb, uint32_t chunk_paths_max, const ymrt_render_settings *rs,
                     int frame, float *noise, uint32_t *instance_ids, const AdjParams *adj, hipStream_t stream,
                     int timing, ymrt_frame_stats *stats);

static int check_settings(const ymrt_render_settings *rs)
{
    YM_REQUIRE(rs->width > 0 && rs->height > 0, "image size must be positive");
    YM_REQUIRE(rs->spp >= 1, "spp must be >= 1");
    YM_REQUIRE(rs->ray_depth >= 1 && rs->ray_depth <= kMaxRayDepth, "ray_depth must be in 1..64");
    YM_REQUIRE(rs->sampler_type == YMRT_SAMPLER_HALTON || rs->sampler_type == YMRT_SAMPLER_PCG,
               "sampler_type must be PCG (0) or HALTON (1)");
    return YMRT_OK;
}

static int ensure_buffers(RenderBuffers *&slot, uint32_t cap)
{
    if (slot && slot->cap >= cap) return YMRT_OK;
    delete slot;
    RenderBuffers *rb = new RenderBuffers;
    slot = rb;
    rb->cap = cap;
    const int rc =
        layout_buffers(*rb, cap, kSpillThreads, [&](auto *&p, size_t count) { return alloc(rb, &p, count); });
    if (rc) {
        delete slot;
        slot = nullptr;
        return rc;
    }
    return YMRT_OK;
}

// Arrays carved out of fixed device memory start on 256-B boundaries.
static size_t carve_round(size_t b) { return (b + 255) & ~(size_t)255; }

size_t fixed_buffer_bytes(uint32_t cap, uint32_t spill_threads)
{
    RenderBuffers sizes;  // owns nothing: only the layout is computed
    size_t total = 0;
    layout_buffers(sizes, cap, spill_threads, [&](auto *&p, size_t count) {
        total += carve_round(sizeof(*p) * std::max<size_t>(count, 1));
        p = nullptr;
        return 0;
    });
    return total;
}

// One frame whose path state is carved out of `memory` (device, `bytes`
// long) instead of allocated: chunks of up to cap paths, persistent kernels
// of up to spill_threads / 64 waves.  Allocates no device memory.
int render_frame_fixed(ymrt_scene *sc, void *memory, size_t bytes, uint32_t cap, uint32_t spill_threads,
                       const ymrt_render_settings *rs, int frame, float *noise, hipStream_t stream,
                       ymrt_frame_stats *stats)
{
    if (int r = check_settings(rs)) return r;
    YM_REQUIRE(memory != nullptr, "no fixed device memory");
    YM_REQUIRE(fixed_buffer_bytes(cap, spill_threads) <= bytes, "fixed device memory too small for the capacity");
    YM_REQUIRE(rs->path_capacity <= cap,
               "path_capacity exceeds the fixed capacity of " + std::to_string(cap) + " paths");
    const uint32_t spp = (uint32_t)rs->spp;
    YM_REQUIRE(spp <= cap, "spp exceeds the fixed capacity of " + std::to_string(cap) + " paths");
    RenderBuffers view;  // owns nothing: no allocations, no events
    view.cap = cap;
    char *base = static_cast<char *>(memory);
    size_t off = 0;
    layout_buffers(view, cap, spill_threads, [&](auto *&p, size_t count) {
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + off);
        off += carve_round(sizeof(*p) * std::max<size_t>(count, 1));
        return 0;
    });
    const uint64_t total_paths = (uint64_t)((rs->width + 15) / 16) * ((rs->height + 15) / 16) * 256ull * spp;
    const uint64_t want = rs->path_capacity ? rs->path_capacity : cap;
    YM_REQUIRE(want >= spp, "path_capacity must hold one pixel's samples");
    const uint64_t per_cap = (want / spp) * spp;
    const uint64_t nchunks = (total_paths + per_cap - 1) / per_cap;
    const uint32_t chunk_paths_max =
        (uint32_t)std::min<uint64_t>(per_cap, (total_paths / nchunks + spp - 1) / spp * spp);
    YM_HIP(hipSetDevice(sc->ctx->device));
    if (int r = clear_stack_flag(sc, stream)) return r;
    return run_frame(sc, &view, chunk_paths_max, rs, frame, noise, nullptr, nullptr, stream, 0, stats);
}

static hipEvent_t next_event(RenderBuffers *rb)
{
    if (rb->event_next == rb->events.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        rb->events.push_back(e);
    }
    return rb->events[rb->event_next++];
}

static int ensure_adjoint(RenderBuffers *rb)
{
    if (rb->adj_slot) return YMRT_OK;
    int rc = alloc(rb, &rb->ps.adj_g, rb->cap);
    rc |= alloc(rb, &rb->adj_counters, kCountersPerSlot);
    rc |= alloc(rb, &rb->adj_stats, kStatCount);
    rc |= alloc(rb, &rb->adj_slot, 1);
    return rc;
}

int render_frame_with(ymrt_scene *sc, RenderBuffers *&buffers, const ymrt_render_settings *rs, int frame,
                      float *noise, uint32_t *instance_ids, const AdjParams *adj, hipStream_t stream, int timing,
                      ymrt_frame_stats *stats)
{
    if (int r = check_settings(rs)) return r;
    YM_HIP(hipSetDevice(sc->ctx->device));
    const uint32_t spp = (uint32_t)rs->spp;
    const uint64_t total_paths = (uint64_t)((rs->width + 15) / 16) * ((rs->height + 15) / 16) * 256ull * spp;
    std::unique_lock<std::mutex> alloc_lock(g_alloc_mu);
    uint64_t cap = rs->path_capacity;
    if (cap == 0) {
        const uint64_t want = std::min<uint64_t>(kDefaultCapacity, total_paths);
        if (buffers && buffers->cap >= want) {
            cap = want;
        } else {
            // grow within a quarter of the free memory (counting the buffers
            // this call would give back), never below what it already holds
            size_t free_b = 0, total_b = 0;
            YM_HIP(hipMemGetInfo(&free_b, &total_b));
            const uint64_t held = buffers ? (uint64_t)buffers->cap * kPathBytes : 0;
            cap = std::min<uint64_t>(want, (free_b + held) / 4 / kPathBytes);
            if (buffers) cap = std::max<uint64_t>(cap, buffers->cap);
        }
        cap = std::max<uint64_t>(cap, spp);
    }
    YM_REQUIRE(cap >= spp, "path_capacity must hold one pixel's samples");
    YM_REQUIRE(rs->persistent_waves <= kMaxPersistentWaves,
               "persistent_waves exceeds " + std::to_string(kMaxPersistentWaves) + " (the traversal spill area)");
    // equal chunks of whole pixels, none larger than the capacity
    const uint64_t per_cap = (cap / spp) * spp;
    const uint64_t nchunks = (total_paths + per_cap - 1) / per_cap;
    const uint32_t chunk_paths_max =
        (uint32_t)std::min<uint64_t>(per_cap, (total_paths / nchunks + spp - 1) / spp * spp);
    int rc = ensure_buffers(buffers, chunk_paths_max);
    if (rc) return rc;
    alloc_lock.unlock();
    return run_frame(sc, buffers, chunk_paths_max, rs, frame, noise, instance_ids, adj, stream, timing, stats);
}

// The frame itself, in chunks of up to chunk_paths_max paths in rb's path
// state (rb->cap >= chunk_paths_max, a multiple of spp).
static int run_frame(ymrt_scene *sc, RenderBuffers *rb, uint32_t chunk_paths_max, const ymrt_render_settings *rs,
                     int frame, float *noise, uint32_t *instance_ids, const AdjParams *adj, hipStream_t stream,
                     int timing, ymrt_frame_stats *stats)
{
    const uint32_t spp = (uint32_t)rs->spp;
    const uint64_t total_paths = (uint64_t)((rs->width + 15) / 16) * ((rs->height + 15) / 16) * 256ull * spp;
    YM_REQUIRE(chunk_paths_max >= spp && chunk_paths_max % spp == 0 && chunk_paths_max <= rb->cap,
               "chunk size must be whole pixels within the path capacity");
    YM_REQUIRE(rs->persistent_waves <= rb->ps.spill_threads / 64u,
               "persistent_waves exceeds " + std::to_string(rb->ps.spill_threads / 64u) +
                   " (the traversal spill area)");
    int rc = YMRT_OK;
    PtParams prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.rs = *rs;
    prm.henum = HaltonEnum::make((int)rs->width, (int)rs->height);
    prm.frame = frame;
    prm.tiles_x = (rs->width + 15) / 16;
    const uint32_t tiles_y = (rs->height + 15) / 16;
    prm.cap = rb->cap;
    prm.ccount_stride = rb->cap / kCompactSpan + 1;
    prm.counters = rb->counters;
    prm.stats = rb->stats;
    prm.error = sc->d_error;
    prm.aov = instance_ids;
    if (adj) {
        YM_REQUIRE(!rs->enable_russian_roulette, "the adjoint needs Russian roulette off");
        YM_REQUIRE(!rs->enable_indirect_clamp, "the adjoint needs the indirect clamp off");
        YM_REQUIRE(rb->allocs.size() > 0, "the adjoint needs allocated path state");
        if ((rc = ensure_adjoint(rb))) return rc;
    }
    YM_REQUIRE(total_paths == (uint64_t)prm.tiles_x * tiles_y * 256ull * spp, "tile count");
    // The coincident-surface list only changes results where medium records
    // change: boundaries, dielectrics, priorities.
    const bool nearby = sc->needs_nearby;
    const bool media = rs->enable_volume_scattering && sc->dev.volume_count > 0;
    prm.media = media ? 1u : 0u;
    const uint32_t stack = sc->max_stack;
    YM_REQUIRE(stack <= (uint32_t)kMaxStackDepth, "scene BVH too deep for the path-tracing traversal stack");
    prm.noise = reinterpret_cast<float4 *>(noise);
    // the frame runs the multi-frame kernels with one slot: the chunk's scene,
    // parameters and path state in a device table of one entry
    const MultiVariant v{nearby, media, sc->dev.has_spheres != 0, sc->dev.has_graph != 0,
                         media ? (sc->dev.has_tex_media ? 2 : sc->dev.grid_count > 0 ? 1 : 0) : 0};
    FrameSlot slot;
    std::memset(&slot, 0, sizeof(slot));
    slot.sc = sc->dev;
    slot.ps = rb->ps;
    const FrameSlot *fs = rb->slots;

    YM_HIP(hipMemsetAsync(rb->stats, 0, sizeof(unsigned long long) * kStatCount, stream));
    rb->event_next = 0;
    struct Mark {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    auto timed = [&](int stage, auto &&fn) -> int {
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            a = next_event(rb);
            b = next_event(rb);
            if (a) (void)hipEventRecord(a, stream);
        }
        int r = fn();
        if (timing && b) {
            (void)hipEventRecord(b, stream);
            marks.push_back(Mark{stage, a, b});
        }
        return r;
    };
    hipEvent_t frame_a = nullptr, frame_b = nullptr;
    if (timing) {
        frame_a = next_event(rb);
        frame_b = next_event(rb);
        YM_HIP(hipEventRecord(frame_a, stream));
    }
    // raygen, the bounce passes and the flush of one chunk over slot table f
    // and its counters; the adjoint's replay runs them again with the
    // event-emitting kernels (timed as one stage)
    auto run_passes = [&](const FrameSlot *f, uint32_t *ctrs, const MultiVariant &vv, bool replay, uint32_t waves,
                          uint32_t blocks, uint32_t &npass) -> int {
        auto st = [&](int stage) { return replay ? 7 : stage; };
        int r = 0;
        if ((r = timed(st(0), [&] { return launch_multi_raygen(f, 1, waves, stream); }))) return r;
        uint32_t pass = 0;
        uint32_t next_check = (uint32_t)rs->ray_depth + 1;
        for (;;) {
            if ((r = timed(st(1), [&] { return launch_multi_extend(f, 1, vv, waves, pass, stream); }))) return r;
            if (media && (r = timed(st(5), [&] {
                              int q = launch_multi_compact(f, 1, blocks, pass, 0, stream);
                              return q ? q : launch_multi_medium(f, 1, vv, waves, pass, stream);
                          })))
                return r;
            if ((r = timed(st(2), [&] { return launch_multi_compact(f, 1, blocks, pass, 1, stream); }))) return r;
            if ((r = timed(st(2), [&] { return launch_multi_shade(f, 1, vv, false, waves, pass, stream); }))) return r;
            if (media && (r = timed(st(2), [&] { return launch_multi_shade(f, 1, vv, true, waves, pass, stream); })))
                return r;
            if ((r = timed(st(3), [&] { return launch_multi_shadow(f, 1, vv, waves, pass, stream); }))) return r;
            if (media && (r = timed(st(6), [&] {
                              int q = launch_multi_compact(f, 1, blocks, pass, 2, stream);
                              return q ? q : launch_multi_transmit(f, 1, vv, waves, pass, stream);
                          })))
                return r;
            ++pass;
            if (pass == next_check || pass >= kMaxPasses) {
                uint32_t active = 0;
                YM_HIP(hipMemcpyAsync(&active, ctrs + pass * kCtrStride + kCtrActive, sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, stream));
                YM_HIP(hipStreamSynchronize(stream));
                if (active == 0) break;
                if (pass >= kMaxPasses) {
                    if ((r = launch_multi_flush(f, 1, pass, stream))) return r;
                    break;
                }
                next_check = pass + 2;
            }
        }
        npass = pass;
        return YMRT_OK;
    };
    MultiVariant vadj = v;
    vadj.adj = true;
    FrameSlot rslot;
    std::memset(&rslot, 0, sizeof(rslot));
    uint32_t passes_total = 0, chunks = 0;
    for (uint64_t base = 0; base < total_paths; base += chunk_paths_max) {
        prm.chunk_base = base;
        prm.chunk_paths = (uint32_t)std::min<uint64_t>(chunk_paths_max, total_paths - base);
        ++chunks;
        // the previous chunk's kernels read the device slot until the stream drains
        if (chunks > 1) YM_HIP(hipStreamSynchronize(stream));
        slot.prm = prm;
        YM_HIP(hipMemcpyAsync(rb->slots, &slot, sizeof(slot), hipMemcpyHostToDevice, stream));
        YM_HIP(hipMemsetAsync(rb->counters, 0, sizeof(uint32_t) * kCountersPerSlot, stream));
        if (adj) {
            // the replay: the same chunk with its own counters and statistics
            rslot = slot;
            rslot.prm.counters = rb->adj_counters;
            rslot.prm.stats = rb->adj_stats;
            rslot.prm.adj = *adj;
            YM_HIP(hipMemcpyAsync(rb->adj_slot, &rslot, sizeof(rslot), hipMemcpyHostToDevice, stream));
            YM_HIP(hipMemsetAsync(rb->adj_counters, 0, sizeof(uint32_t) * kCountersPerSlot, stream));
        }
        const uint32_t waves = persistent_waves(prm.chunk_paths, rs->persistent_waves, rb->ps.spill_threads);
        const uint32_t blocks = (prm.chunk_paths + kCompactSpan - 1) / kCompactSpan;
        uint32_t npass = 0;
        if ((rc = run_passes(fs, rb->counters, v, false, waves, blocks, npass))) return rc;
        passes_total += npass;
        const uint32_t pixel_blocks = (prm.chunk_paths / spp + 255) / 256;
        if ((rc = timed(4, [&] { return launch_multi_resolve(fs, 1, pixel_blocks, stream); }))) return rc;
        if (adj && (rc = run_passes(rb->adj_slot, rb->adj_counters, vadj, true, waves, blocks, npass))) return rc;
    }
    if (timing) YM_HIP(hipEventRecord(frame_b, stream));
    if (stats) {
        unsigned long long st[kStatCount];
        YM_HIP(hipMemcpyAsync(st, rb->stats, sizeof(st), hipMemcpyDeviceToHost, stream));
        YM_HIP(hipStreamSynchronize(stream));
        std::memset(stats, 0, sizeof(*stats));
        stats->camera_rays = st[kStatCamera];
        stats->extension_rays = st[kStatExtension];
        stats->shadow_rays = st[kStatShadow];
        stats->transmittance_rays = st[kStatTransmittance];
        stats->extend_nodes = st[kStatExtendNodes];
        stats->extend_tris = st[kStatExtendTris];
        stats->shadow_nodes = st[kStatShadowNodes];
        stats->shadow_tris = st[kStatShadowTris];
        stats->transmit_nodes = st[kStatTransmitNodes];
        stats->transmit_tris = st[kStatTransmitTris];
        stats->paths = (uint64_t)rs->width * rs->height * spp;
        stats->passes = passes_total;
        stats->chunks = chunks;
        if (timing) {
            float ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            uint32_t launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (const Mark &m : marks) {
                float t = 0.0f;
                YM_HIP(hipEventElapsedTime(&t, m.a, m.b));
                ms[m.stage] += t;
                launches[m.stage]++;
            }
            stats->ms_raygen = ms[0];
            stats->ms_extend = ms[1];
            stats->ms_shade = ms[2];
            stats->ms_shadow = ms[3];
            stats->ms_resolve = ms[4];
            stats->ms_medium = ms[5];
            stats->launches_medium = launches[5];
            stats->ms_transmit = ms[6];
            stats->launches_transmit = launches[6];
            stats->ms_adjoint = ms[7];
            stats->launches_extend = launches[1];
            stats->launches_shade = launches[2];
            stats->launches_shadow = launches[3];
            float tot = 0.0f;
            YM_HIP(hipEventElapsedTime(&tot, frame_a, frame_b));
            stats->ms_total = tot;
        }
    } else {
        YM_HIP(hipStreamSynchronize(stream));
    }
    return YMRT_OK;
}

int render_frame(ymrt_scene *sc, const ymrt_render_settings *rs, int frame, float *noise, uint32_t *instance_ids,
                 const AdjParams *adj, hipStream_t stream, int timing, ymrt_frame_stats *stats)
{
    YM_HIP(hipSetDevice(sc->ctx->device));
    const int rc = clear_stack_flag(sc, stream);
    if (rc) return rc;
    return render_frame_with(sc, sc->render, rs, frame, noise, instance_ids, adj, stream, timing, stats);
}

}  // namespace ym

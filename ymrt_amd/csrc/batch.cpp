// batch.cpp -- many frames, of one or many scenes, in shared launches.
//
// YMRT renders each view from its own render thread and stream, one frame
// after another (src/core/RenderModule.cpp:55-66 and the render threads at
// 95-143), and a frame's host loop launches every stage and reads counters
// back once per stage (src/core/RenderModule.cpp:580-631).  Small frames then
// leave most of the GPU idle: a few thousand paths per stage, launch latency
// and host round trips between them, and each frame's tail of long paths.
//
// Here the frames of a batch are cut into chunks of whole pixels and packed
// into rounds: a round is up to kMaxSlots chunks (slots) whose paths share
// one path-state allocation, each slot in its own range of it, with its own
// pass counters and queues.  One launch per stage serves every slot of the
// round (the *_m kernels walk a device table of slots: each wave starts at
// the slot its index falls in by share of work and helps the others when
// that one is drained), and the host checks for remaining paths once per
// round after the longest ray depth, as the single-frame loop does per chunk.
// Small frames therefore fill the GPU together and share one tail.  Frames do
// not depend on how they are chunked or packed (each path writes its own
// output slot, the resolve sums samples in a fixed order), so every frame is
// bit-identical to ymrt_render_frame.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "buffers.hpp"
#include "internal.hpp"
#include "pt.hpp"

namespace ym {

uint32_t persistent_waves(uint64_t paths, uint32_t requested, uint32_t spill_threads);
int launch_multi_active(const FrameSlot *fs, uint32_t nfs, uint32_t pass, uint32_t *out, hipStream_t s);
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

// Same sizing rule as a single frame (render.cpp): up to 64 M paths in
// flight, within a quarter of the free device memory.
constexpr uint64_t kBatchCapacity = 1ull << 26;
constexpr uint64_t kBatchPathBytes = 720;

namespace {

struct Chunk {
    uint32_t item;
    uint64_t base;
    uint32_t paths;
};

struct Round {
    std::vector<Chunk> slots;
    uint64_t used = 0;
};

uint64_t frame_paths(const ymrt_render_settings &rs)
{
    return (uint64_t)((rs.width + 15) / 16) * ((rs.height + 15) / 16) * 256ull * (uint64_t)rs.spp;
}

int het_of(const DevScene &sc) { return sc.has_tex_media ? 2 : sc.grid_count > 0 ? 1 : 0; }

}  // namespace

// A context's batch buffers: the shared path state of a round (grown as
// batches need, kept for the next call, released with the context), the slot
// table and the word the host reads to end a round.
struct BatchMemory {
    RenderBuffers *rb = nullptr;  // path state, counters, statistics, slot table
    uint32_t stat_sets = 0;
    FrameSlot *h_table = nullptr; // pinned host copy of the slot table
    uint32_t *d_active = nullptr;
    ~BatchMemory()
    {
        delete rb;
        if (h_table) (void)hipHostFree(h_table);
        if (d_active) (void)hipFree(d_active);
    }
};

void free_batch_memory(BatchMemory *m) { delete m; }

static int ensure_batch_memory(ymrt_context *ctx, uint64_t cap, uint32_t frames)
{
    if (!ctx->batch) ctx->batch = new BatchMemory;
    BatchMemory &m = *ctx->batch;
    if (!m.h_table) {
        YM_HIP(hipHostMalloc(reinterpret_cast<void **>(&m.h_table), sizeof(FrameSlot) * kMaxSlots));
        YM_HIP(hipMalloc(&m.d_active, sizeof(uint32_t)));
    }
    if (m.rb && m.rb->cap >= cap && m.stat_sets >= frames) return YMRT_OK;
    delete m.rb;
    m.rb = nullptr;
    RenderBuffers *rb = new RenderBuffers;
    rb->cap = (uint32_t)cap;
    const int rc = layout_buffers(
        *rb, cap, kSpillThreads, [&](auto *&p, size_t count) { return alloc(rb, &p, count); }, kMaxSlots, frames);
    if (rc) {
        delete rb;
        return rc;
    }
    m.rb = rb;
    m.stat_sets = frames;
    return YMRT_OK;
}

int render_batch(ymrt_context *ctx, const ymrt_batch_item *items, uint32_t n, uint32_t max_slots,
                 hipStream_t stream)
{
    YM_REQUIRE(ctx != nullptr, "context is NULL");
    YM_REQUIRE(n == 0 || items != nullptr, "items is NULL");
    YM_REQUIRE(max_slots <= kMaxSlots, "max_slots exceeds " + std::to_string(kMaxSlots));
    for (uint32_t i = 0; i < n; ++i) {
        const ymrt_batch_item &it = items[i];
        const std::string at = "batch item " + std::to_string(i) + ": ";
        YM_REQUIRE(it.scene != nullptr && it.settings != nullptr && it.noise_rgba != nullptr,
                   at + "NULL scene, settings or image");
        YM_REQUIRE(it.scene->ctx == ctx, at + "scene of another context");
        YM_REQUIRE(it.frame_index >= 1, at + "frame_index starts at 1");
        const ymrt_render_settings &rs = *it.settings;
        YM_REQUIRE(rs.width > 0 && rs.height > 0, at + "image size must be positive");
        YM_REQUIRE(rs.spp >= 1, at + "spp must be >= 1");
        YM_REQUIRE(rs.ray_depth >= 1 && rs.ray_depth <= kMaxRayDepth, at + "ray_depth must be in 1..64");
        YM_REQUIRE(rs.sampler_type == YMRT_SAMPLER_HALTON || rs.sampler_type == YMRT_SAMPLER_PCG,
                   at + "sampler_type must be PCG (0) or HALTON (1)");
        YM_REQUIRE(it.scene->max_stack <= (uint32_t)kMaxStackDepth,
                   at + "scene BVH too deep for the path-tracing traversal stack");
    }
    YM_HIP(hipSetDevice(ctx->device));
    if (n == 0) return YMRT_OK;
    const uint32_t slots_max = max_slots ? max_slots : kMaxSlots;

    // every scene's overflow flag covers all of its frames in this call
    for (uint32_t i = 0; i < n; ++i)
        if (int rc = clear_stack_flag(items[i].scene, stream)) return rc;

    // one batch at a time per context: the round's buffers are the context's
    std::lock_guard<std::mutex> lock(ctx->batch_mu);
    // capacity of a round: what the batch needs, within a quarter of the free
    // memory (counting the context's buffers), never below what it holds
    uint64_t total = 0, spp_max = 1;
    for (uint32_t i = 0; i < n; ++i) {
        total += frame_paths(*items[i].settings);
        spp_max = std::max<uint64_t>(spp_max, (uint64_t)items[i].settings->spp);
    }
    const uint64_t held = ctx->batch && ctx->batch->rb ? ctx->batch->rb->cap : 0;
    uint64_t cap = std::min<uint64_t>(kBatchCapacity, total);
    if (cap > held) {
        size_t free_b = 0, total_b = 0;
        YM_HIP(hipMemGetInfo(&free_b, &total_b));
        cap = std::min<uint64_t>(cap, ((uint64_t)free_b + held * kBatchPathBytes) / 4 / kBatchPathBytes);
    }
    cap = std::max<uint64_t>(std::max<uint64_t>(cap, spp_max), std::min<uint64_t>(held, kBatchCapacity));
    YM_REQUIRE(cap <= 0xFFFFFFFFull, "batch capacity overflow");

    // chunks of whole pixels, none larger than a round, equal per frame
    std::vector<Chunk> chunks;
    std::vector<uint32_t> chunks_of(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t spp = (uint64_t)items[i].settings->spp, tp = frame_paths(*items[i].settings);
        const uint64_t per_cap = (cap / spp) * spp;
        const uint64_t nch = (tp + per_cap - 1) / per_cap;
        const uint64_t size = std::min<uint64_t>(per_cap, (tp / nch + spp - 1) / spp * spp);
        for (uint64_t b = 0; b < tp; b += size) {
            chunks.push_back(Chunk{i, b, (uint32_t)std::min<uint64_t>(size, tp - b)});
            chunks_of[i]++;
        }
    }
    // rounds: first fit, largest chunks first
    std::stable_sort(chunks.begin(), chunks.end(), [](const Chunk &a, const Chunk &b) { return a.paths > b.paths; });
    std::vector<Round> rounds;
    for (const Chunk &c : chunks) {
        Round *r = nullptr;
        for (Round &q : rounds)
            if (q.used + c.paths <= cap && q.slots.size() < slots_max) {
                r = &q;
                break;
            }
        if (!r) {
            rounds.emplace_back();
            r = &rounds.back();
        }
        r->slots.push_back(c);
        r->used += c.paths;
    }

    if (int rc = ensure_batch_memory(ctx, cap, n)) return rc;
    BatchMemory &mem = *ctx->batch;
    RenderBuffers &rb = *mem.rb;
    YM_HIP(hipMemsetAsync(rb.stats, 0, sizeof(unsigned long long) * kStatCount * n, stream));
    const uint32_t ccount_stride = rb.cap / kCompactSpan + 1;
    std::vector<uint32_t> passes_of(n, 0);

    for (const Round &r : rounds) {
        const uint32_t nfs = (uint32_t)r.slots.size();
        MultiVariant v{false, false, false, false, 0};
        uint64_t off = 0, paths = 0;
        uint32_t depth_max = 1, compact_blocks = 0, pixel_blocks = 0;
        for (uint32_t s = 0; s < nfs; ++s) {
            const Chunk &c = r.slots[s];
            const ymrt_batch_item &it = items[c.item];
            const ymrt_render_settings &rs = *it.settings;
            ymrt_scene *sc = it.scene;
            FrameSlot &f = mem.h_table[s];
            std::memset(&f, 0, sizeof(f));
            f.sc = sc->dev;
            PtParams &prm = f.prm;
            prm.rs = rs;
            prm.henum = HaltonEnum::make((int)rs.width, (int)rs.height);
            prm.frame = it.frame_index;
            prm.tiles_x = (rs.width + 15) / 16;
            prm.chunk_base = c.base;
            prm.chunk_paths = c.paths;
            prm.cap = rb.cap;
            prm.ccount_stride = ccount_stride;
            prm.counters = rb.counters + (size_t)s * kCountersPerSlot;
            prm.stats = rb.stats + (size_t)c.item * kStatCount;
            prm.error = sc->d_error;
            prm.noise = reinterpret_cast<float4 *>(it.noise_rgba);
            const bool media = rs.enable_volume_scattering && sc->dev.volume_count > 0;
            prm.media = media ? 1u : 0u;
            // the slot's range of the shared path state (per-path arrays and
            // queues by offset; arrays of several entries per path keep the
            // round's capacity as their stride)
            PathState ps = rb.ps;
            auto sh = [&](auto *&p) { p += off; };
            sh(ps.ray_o), sh(ps.ray_d), sh(ps.last), sh(ps.beta), sh(ps.L), sh(ps.smp), sh(ps.info), sh(ps.xhdr);
            sh(ps.xrec), sh(ps.hit_a), sh(ps.hit_b), sh(ps.nb), sh(ps.sh_o), sh(ps.sh_d), sh(ps.sh_f), sh(ps.sh_li);
            sh(ps.sh_tr), sh(ps.sh_c), sh(ps.sh_flags), sh(ps.out), sh(ps.queue[0]), sh(ps.queue[1]), sh(ps.q_hit);
            sh(ps.q_hit_vol), sh(ps.q_hit_diffuse), sh(ps.q_hit_metal), sh(ps.q_medium), sh(ps.q_shadow);
            sh(ps.q_transmit), sh(ps.flag), sh(ps.first_hit);
            ps.ccount = rb.ps.ccount + (size_t)s * 4 * ccount_stride;
            f.ps = ps;
            off += c.paths;
            paths += c.paths;
            v.nearby = v.nearby || sc->needs_nearby;
            v.media = v.media || media;
            v.sphere = v.sphere || sc->dev.has_spheres;
            v.graph = v.graph || sc->dev.has_graph;
            if (media) v.het = std::max(v.het, het_of(sc->dev));
            depth_max = std::max<uint32_t>(depth_max, (uint32_t)rs.ray_depth);
            compact_blocks = std::max<uint32_t>(compact_blocks, (c.paths + kCompactSpan - 1) / kCompactSpan);
            pixel_blocks = std::max<uint32_t>(pixel_blocks, (c.paths / (uint32_t)rs.spp + 255) / 256);
        }
        YM_HIP(hipMemcpyAsync(rb.slots, mem.h_table, sizeof(FrameSlot) * nfs, hipMemcpyHostToDevice, stream));
        YM_HIP(hipMemsetAsync(rb.counters, 0, sizeof(uint32_t) * kCountersPerSlot * nfs, stream));
        // one wave per 256 paths of the round, within [1024, the spill area]
        const uint32_t waves = persistent_waves(paths, 0, rb.ps.spill_threads);
        const FrameSlot *fs = rb.slots;
        int rc;
        if ((rc = launch_multi_raygen(fs, nfs, waves, stream))) return rc;
        uint32_t pass = 0, next_check = depth_max + 1;
        for (;;) {
            if ((rc = launch_multi_extend(fs, nfs, v, waves, pass, stream))) return rc;
            if (v.media) {
                if ((rc = launch_multi_compact(fs, nfs, compact_blocks, pass, 0, stream))) return rc;
                if ((rc = launch_multi_medium(fs, nfs, v, waves, pass, stream))) return rc;
            }
            if ((rc = launch_multi_compact(fs, nfs, compact_blocks, pass, 1, stream))) return rc;
            if ((rc = launch_multi_shade(fs, nfs, v, false, waves, pass, stream))) return rc;
            if (v.media && (rc = launch_multi_shade(fs, nfs, v, true, waves, pass, stream))) return rc;
            if ((rc = launch_multi_shadow(fs, nfs, v, waves, pass, stream))) return rc;
            if (v.media) {
                if ((rc = launch_multi_compact(fs, nfs, compact_blocks, pass, 2, stream))) return rc;
                if ((rc = launch_multi_transmit(fs, nfs, v, waves, pass, stream))) return rc;
            }
            ++pass;
            if (pass == next_check || pass >= kMaxPasses) {
                uint32_t active = 0;
                if ((rc = launch_multi_active(fs, nfs, pass, mem.d_active, stream))) return rc;
                YM_HIP(hipMemcpyAsync(&active, mem.d_active, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
                YM_HIP(hipStreamSynchronize(stream));
                if (active == 0) break;
                if (pass >= kMaxPasses) {
                    if ((rc = launch_multi_flush(fs, nfs, pass, stream))) return rc;
                    break;
                }
                next_check = pass + 2;
            }
        }
        if ((rc = launch_multi_resolve(fs, nfs, pixel_blocks, stream))) return rc;
        for (const Chunk &c : r.slots) passes_of[c.item] += pass;
        // the host table is rewritten for the next round
        YM_HIP(hipStreamSynchronize(stream));
    }

    std::vector<unsigned long long> st((size_t)kStatCount * n);
    YM_HIP(hipMemcpyAsync(st.data(), rb.stats, sizeof(unsigned long long) * st.size(), hipMemcpyDeviceToHost, stream));
    YM_HIP(hipStreamSynchronize(stream));
    for (uint32_t i = 0; i < n; ++i) {
        ymrt_frame_stats *out = items[i].stats;
        if (!out) continue;
        const unsigned long long *s = st.data() + (size_t)i * kStatCount;
        std::memset(out, 0, sizeof(*out));
        out->camera_rays = s[kStatCamera];
        out->extension_rays = s[kStatExtension];
        out->shadow_rays = s[kStatShadow];
        out->transmittance_rays = s[kStatTransmittance];
        out->extend_nodes = s[kStatExtendNodes];
        out->extend_tris = s[kStatExtendTris];
        out->shadow_nodes = s[kStatShadowNodes];
        out->shadow_tris = s[kStatShadowTris];
        out->transmit_nodes = s[kStatTransmitNodes];
        out->transmit_tris = s[kStatTransmitTris];
        const ymrt_render_settings &rs = *items[i].settings;
        out->paths = (uint64_t)rs.width * rs.height * (uint64_t)rs.spp;
        out->passes = passes_of[i];
        out->chunks = chunks_of[i];
    }
    // a traversal overflow in any frame fails the batch, as it fails a frame
    for (uint32_t i = 0; i < n; ++i) {
        const int rc = check_stack_flag(items[i].scene);
        if (rc) {
            set_error("batch item " + std::to_string(i) + ": " + ymrt_last_error());
            return rc;
        }
    }
    return YMRT_OK;
}

}  // namespace ym

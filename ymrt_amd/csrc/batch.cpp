// batch.cpp -- several frames, of one or many scenes, in one call.
//
// YMRT renders its editor view and its path-tracing view from two host
// threads, each driving its own stream (src/core/RenderModule.cpp:55-66 and
// the render threads at 95-143).  A batch does the same for any number of
// frames: a few host threads, each with a stream of its own (default
// priority), take scenes longest first and render their frames with the
// single-frame path, so the small launches and per-chunk counter reads of one
// frame overlap with another frame's kernels.  Frames of one scene run in
// item order on one stream.  Each thread owns one set of path-state buffers
// for every frame it renders (grown as frames need, freed at the end), so
// the memory a batch holds does not grow with its scene count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace ym {

constexpr uint32_t kDefaultBatchStreams = 4;

// Work estimate of one frame: samples x bounces x the depth of a tree over
// the scene's triangles.
static double frame_cost(const ymrt_batch_item &it)
{
    const ymrt_render_settings &rs = *it.settings;
    const double tris = (double)it.scene->info.triangle_count;
    return (double)rs.width * rs.height * std::max(rs.spp, 1) * std::max(rs.ray_depth, 1) * std::log2(2.0 + tris);
}

int render_batch(ymrt_context *ctx, const ymrt_batch_item *items, uint32_t n, uint32_t max_streams,
                 hipStream_t stream)
{
    YM_REQUIRE(ctx != nullptr, "context is NULL");
    YM_REQUIRE(n == 0 || items != nullptr, "items is NULL");
    for (uint32_t i = 0; i < n; ++i) {
        const ymrt_batch_item &it = items[i];
        YM_REQUIRE(it.scene != nullptr && it.settings != nullptr && it.noise_rgba != nullptr,
                   "batch item " + std::to_string(i) + ": NULL scene, settings or image");
        YM_REQUIRE(it.scene->ctx == ctx, "batch item " + std::to_string(i) + ": scene of another context");
        YM_REQUIRE(it.frame_index >= 1, "batch item " + std::to_string(i) + ": frame_index starts at 1");
    }
    // the caller's earlier work on stream (images, settings) comes first
    YM_HIP(hipStreamSynchronize(stream));
    if (n == 0) return YMRT_OK;

    // one group per scene, its frames in item order; groups longest first
    std::vector<std::vector<uint32_t>> groups;
    std::unordered_map<const ymrt_scene *, size_t> group_of;
    std::vector<double> cost;
    for (uint32_t i = 0; i < n; ++i) {
        auto f = group_of.find(items[i].scene);
        size_t g;
        if (f == group_of.end()) {
            g = groups.size();
            group_of.emplace(items[i].scene, g);
            groups.emplace_back();
            cost.push_back(0.0);
        } else {
            g = f->second;
        }
        groups[g].push_back(i);
        cost[g] += frame_cost(items[i]);
    }
    std::vector<size_t> order(groups.size());
    for (size_t g = 0; g < order.size(); ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });

    const uint32_t streams =
        std::min<uint32_t>((uint32_t)groups.size(), max_streams ? max_streams : kDefaultBatchStreams);
    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex err_mu;
    int err_rc = YMRT_OK;
    std::string err_msg;
    auto fail = [&](int rc, const std::string &msg) {
        std::lock_guard<std::mutex> lk(err_mu);
        if (err_rc == YMRT_OK) {
            err_rc = rc;
            err_msg = msg;
        }
        failed = true;
    };
    auto worker = [&]() {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            fail(YMRT_ERR_DEVICE, "hipSetDevice failed in a batch thread");
            return;
        }
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            fail(YMRT_ERR_DEVICE, "hipStreamCreateWithFlags failed in a batch thread");
            return;
        }
        RenderBuffers *buffers = nullptr;
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= order.size() || failed) break;
            // the scene's overflow flag covers all of its frames in this batch
            if (int rc = clear_stack_flag(items[groups[order[k]].front()].scene, s)) {
                fail(rc, std::string("batch: ") + ymrt_last_error());
                break;
            }
            for (uint32_t i : groups[order[k]]) {
                const ymrt_batch_item &it = items[i];
                int rc;
                try {
                    rc = render_frame_with(it.scene, buffers, it.settings, it.frame_index, it.noise_rgba, nullptr,
                                           nullptr, nullptr, s, it.timing, it.stats);
                } catch (const std::exception &e) {
                    set_error(std::string("render failed: ") + e.what());
                    rc = YMRT_ERR_INTERNAL;
                }
                if (rc) {
                    fail(rc, "batch item " + std::to_string(i) + ": " + ymrt_last_error());
                    break;
                }
            }
        }
        if (hipStreamSynchronize(s) != hipSuccess) fail(YMRT_ERR_DEVICE, "batch stream failed");
        free_render_buffers(buffers);
        if (hipStreamDestroy(s) != hipSuccess) fail(YMRT_ERR_DEVICE, "hipStreamDestroy failed in a batch thread");
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 1; w < streams; ++w) pool.emplace_back(worker);
    worker();
    for (std::thread &t : pool) t.join();
    if (err_rc != YMRT_OK) {
        set_error(err_msg);
        return err_rc;
    }
    // a traversal overflow in any scene's frames fails the batch, as it fails
    // a single frame
    for (size_t g = 0; g < groups.size(); ++g) {
        const uint32_t i = groups[g].front();
        const int rc = check_stack_flag(items[i].scene);
        if (rc) {
            set_error("batch item " + std::to_string(i) + ": " + ymrt_last_error());
            return rc;
        }
    }
    return YMRT_OK;
}

}  // namespace ym

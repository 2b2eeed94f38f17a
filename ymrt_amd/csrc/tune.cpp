// tune.cpp -- timed trials of the launch settings that change no pixel.
//
// A frame's chunk size (path_capacity) trades the idle tail passes of each
// chunk, which hold only its longest paths, against memory; the wave count of
// the persistent pass kernels (persistent_waves) trades occupancy against the
// tail of their dynamic work fetch.  Neither changes a pixel: path state is
// indexed by slot, every queue consumer is order-free and no float is summed
// with atomics, so a tuned setting keeps the image bit for bit.
#include <algorithm>
#include <string>
#include <vector>

#include "internal.hpp"
#include "pt.hpp"

namespace ym {

constexpr int kTrialRepeats = 2;

int autotune(ymrt_scene *sc, const ymrt_render_settings *rs_in, int frame, hipStream_t stream,
             ymrt_tune_result *out)
{
    YM_REQUIRE(sc != nullptr && rs_in != nullptr && out != nullptr, "NULL argument");
    YM_REQUIRE(frame >= 1, "frame_index starts at 1");
    YM_REQUIRE(rs_in->spp >= 1 && rs_in->width >= 1 && rs_in->height >= 1, "empty frame");
    const uint64_t spp = (uint64_t)rs_in->spp;
    const uint64_t total = (uint64_t)((rs_in->width + 15) / 16) * ((rs_in->height + 15) / 16) * 256ull * spp;
    YM_HIP(hipSetDevice(sc->ctx->device));
    // an overflow in any trial frame fails the tuning
    if (int r = clear_stack_flag(sc, stream)) return r;
    float *img = nullptr;
    YM_HIP(hipMalloc(&img, sizeof(float) * 4 * rs_in->width * rs_in->height));
    ymrt_tune_result res{};
    int rc = YMRT_OK;
    // the faster of kTrialRepeats frames; a capacity the device cannot hold
    // drops out (ms < 0) instead of failing the tuning
    auto trial = [&](uint32_t cap, uint32_t waves, float &ms) -> int {
        ymrt_render_settings rs = *rs_in;
        rs.path_capacity = cap;
        rs.persistent_waves = waves;
        ms = -1.0f;
        for (int k = 0; k < kTrialRepeats; ++k) {
            ymrt_frame_stats st{};
            const int r = render_frame_with(sc, sc->render, &rs, frame, img, nullptr, nullptr, stream, 1, &st);
            ++res.trials;
            if (r == YMRT_ERR_OUT_OF_MEMORY && cap != 0) return YMRT_OK;
            if (r) return r;
            if (ms < 0.0f || st.ms_total < ms) ms = st.ms_total;
        }
        return YMRT_OK;
    };
    try {
        ymrt_render_settings warm = *rs_in;
        warm.path_capacity = 0;
        warm.persistent_waves = 0;
        rc = render_frame_with(sc, sc->render, &warm, frame, img, nullptr, nullptr, stream, 0, nullptr);
        ++res.trials;
        std::vector<uint32_t> caps{0};
        for (uint64_t div : {2ull, 4ull}) {
            const uint64_t c = (total / div + spp - 1) / spp * spp;
            if (c >= spp && c < total && c <= 0xFFFFFFFFull) caps.push_back((uint32_t)c);
        }
        float best = -1.0f;
        for (size_t i = 0; !rc && i < caps.size(); ++i) {
            float ms;
            rc = trial(caps[i], 0, ms);
            if (!rc && i == 0) res.ms_default = ms;
            if (!rc && ms >= 0.0f && (best < 0.0f || ms < best)) {
                best = ms;
                res.path_capacity = caps[i];
            }
        }
        for (uint32_t w : {4096u, 2048u, 1024u}) {
            if (rc) break;
            float ms;
            rc = trial(res.path_capacity, w, ms);
            if (!rc && ms >= 0.0f && ms < best) {
                best = ms;
                res.persistent_waves = w;
            }
        }
        res.ms = best;
    } catch (const std::exception &e) {
        set_error(std::string("autotune failed: ") + e.what());
        rc = YMRT_ERR_INTERNAL;
    }
    const hipError_t fe = hipFree(img);
    if (rc) return rc;
    YM_REQUIRE(fe == hipSuccess, "hipFree failed after tuning");
    *out = res;
    return YMRT_OK;
}

}  // namespace ym

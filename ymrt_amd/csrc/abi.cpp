// abi.cpp -- the extern "C" entry points of include/ymrt.h.
#include <cstring>
#include <new>

#include "internal.hpp"
#include "pt.hpp"

namespace ym {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int build_scene(ymrt_scene *sc, const ymrt_scene_desc *d);                       // scene.cpp
int update_materials(ymrt_scene *sc, uint32_t first, uint32_t count, const ymrt_material_desc *m,
                     hipStream_t s);  // scene.cpp
int update_texture(ymrt_scene *sc, uint32_t tex, const float *texels, hipStream_t s);  // scene.cpp
int update_grid(ymrt_scene *sc, uint32_t grid, const float *density, hipStream_t s);    // scene.cpp
int render_batch(ymrt_context *ctx, const ymrt_batch_item *items, uint32_t n, uint32_t max_slots,
                 hipStream_t stream);  // batch.cpp
int autotune(ymrt_scene *sc, const ymrt_render_settings *rs, int frame, hipStream_t stream,
             ymrt_tune_result *out);  // tune.cpp
int render_frame(ymrt_scene *sc, const ymrt_render_settings *rs, int frame, float *noise, uint32_t *instance_ids,
                 const AdjParams *adj, hipStream_t stream, int timing, ymrt_frame_stats *stats);  // render.cpp
int noalloc_setup(int device);  // noalloc.hip
int render_frame_noalloc(ymrt_scene *sc, const ymrt_render_settings *rs, int frame, float *noise,
                         hipStream_t stream, ymrt_frame_stats *stats);  // noalloc.hip
int accumulate(const float *noise, float *accum, uint32_t w, uint32_t h, int frame, hipStream_t s);
int tonemap(float *rgba, uint32_t w, uint32_t h, float gamma, float exposure, hipStream_t s);

// The overflow flag describes one call: cleared on the call's stream before
// its kernels, read after the stream has drained.
int clear_stack_flag(ymrt_scene *sc, hipStream_t stream)
{
    YM_HIP(hipMemsetAsync(sc->d_error, 0, sizeof(uint32_t), stream));
    return YMRT_OK;
}

int check_stack_flag(ymrt_scene *sc)
{
    uint32_t flag = 0;
    YM_HIP(hipMemcpy(&flag, sc->d_error, sizeof(flag), hipMemcpyDeviceToHost));
    if (flag) {
        set_error("traversal stack overflow");
        return YMRT_ERR_INTERNAL;
    }
    return YMRT_OK;
}

}  // namespace ym

using namespace ym;

extern "C" {

const char *ymrt_last_error(void) { return g_last_error.c_str(); }

int ymrt_abi_version(void) { return YMRT_ABI_VERSION; }

int ymrt_context_create(int device, ymrt_context **out)
{
    YM_REQUIRE(out != nullptr, "out is NULL");
    int count = 0;
    YM_HIP(hipGetDeviceCount(&count));
    YM_REQUIRE(device >= 0 && device < count, "no such device");
    YM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    YM_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("built for gfx950, device is ") + prop.gcnArchName);
        return YMRT_ERR_UNSUPPORTED;
    }
    ymrt_context *c = new (std::nothrow) ymrt_context;
    if (!c) {
        set_error("out of host memory");
        return YMRT_ERR_OUT_OF_MEMORY;
    }
    c->device = device;
    // the fixed device memory of ymrt_render_frame_noalloc, once per device
    if (int rc = noalloc_setup(device)) {
        delete c;
        return rc;
    }
    *out = c;
    return YMRT_OK;
}

int ymrt_context_destroy(ymrt_context *ctx)
{
    delete ctx;
    return YMRT_OK;
}

int ymrt_scene_create(ymrt_context *ctx, const ymrt_scene_desc *desc, ymrt_scene **out)
{
    YM_REQUIRE(ctx != nullptr && out != nullptr, "NULL argument");
    YM_HIP(hipSetDevice(ctx->device));
    ymrt_scene *sc = new (std::nothrow) ymrt_scene;
    if (!sc) {
        set_error("out of host memory");
        return YMRT_ERR_OUT_OF_MEMORY;
    }
    sc->ctx = ctx;
    int rc;
    try {
        rc = build_scene(sc, desc);
    } catch (const std::exception &e) {
        set_error(std::string("scene build failed: ") + e.what());
        rc = YMRT_ERR_INTERNAL;
    }
    if (rc != YMRT_OK) {
        ymrt_scene_destroy(sc);
        return rc;
    }
    *out = sc;
    return YMRT_OK;
}

int ymrt_scene_destroy(ymrt_scene *sc)
{
    if (!sc) return YMRT_OK;
    destroy_render_buffers(sc);
    for (void *p : sc->allocations) (void)hipFree(p);
    delete sc;
    return YMRT_OK;
}

int ymrt_scene_release_buffers(ymrt_scene *sc)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_HIP(hipSetDevice(sc->ctx->device));
    YM_HIP(hipDeviceSynchronize());
    destroy_render_buffers(sc);
    return YMRT_OK;
}

int ymrt_scene_get_info(const ymrt_scene *sc, ymrt_scene_info *info)
{
    YM_REQUIRE(sc != nullptr && info != nullptr, "NULL argument");
    *info = sc->info;
    return YMRT_OK;
}

int ymrt_trace_closest(ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_REQUIRE(n == 0 || (rays && hits), "NULL buffer");
    int rc = clear_stack_flag(sc, (hipStream_t)stream);
    if (rc) return rc;
    rc = launch_trace(sc, rays, n, hits, nullptr, nullptr, 0, (hipStream_t)stream);
    if (rc) return rc;
    YM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return check_stack_flag(sc);
}

int ymrt_trace_any(ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, uint32_t *occluded, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_REQUIRE(n == 0 || (rays && occluded), "NULL buffer");
    int rc = clear_stack_flag(sc, (hipStream_t)stream);
    if (rc) return rc;
    rc = launch_trace(sc, rays, n, nullptr, occluded, nullptr, 1, (hipStream_t)stream);
    if (rc) return rc;
    YM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return check_stack_flag(sc);
}

int ymrt_trace_closest_counted(ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits,
                               uint32_t *visits, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_REQUIRE(n == 0 || (rays && hits && visits), "NULL buffer");
    int rc = clear_stack_flag(sc, (hipStream_t)stream);
    if (rc) return rc;
    rc = launch_trace(sc, rays, n, hits, nullptr, visits, 0, (hipStream_t)stream);
    if (rc) return rc;
    YM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return check_stack_flag(sc);
}

int ymrt_render_frame(ymrt_scene *sc, const ymrt_render_settings *rs, int frame_index, float *noise_rgba,
                      void *stream, int timing, ymrt_frame_stats *stats)
{
    YM_REQUIRE(sc != nullptr && rs != nullptr && noise_rgba != nullptr, "NULL argument");
    YM_REQUIRE(frame_index >= 1, "frame_index starts at 1");
    try {
        int rc = render_frame(sc, rs, frame_index, noise_rgba, nullptr, nullptr, (hipStream_t)stream, timing,
                              stats);
        if (rc) return rc;
    } catch (const std::exception &e) {
        set_error(std::string("render failed: ") + e.what());
        return YMRT_ERR_INTERNAL;
    }
    return check_stack_flag(sc);
}

int ymrt_render_frame_noalloc(ymrt_scene *sc, const ymrt_render_settings *rs, int frame_index, float *noise_rgba,
                              void *stream, ymrt_frame_stats *stats)
{
    YM_REQUIRE(sc != nullptr && rs != nullptr && noise_rgba != nullptr, "NULL argument");
    YM_REQUIRE(frame_index >= 1, "frame_index starts at 1");
    try {
        int rc = render_frame_noalloc(sc, rs, frame_index, noise_rgba, (hipStream_t)stream, stats);
        if (rc) return rc;
    } catch (const std::exception &e) {
        set_error(std::string("render failed: ") + e.what());
        return YMRT_ERR_INTERNAL;
    }
    return check_stack_flag(sc);
}

int ymrt_render_adjoint(ymrt_scene *sc, const ymrt_render_settings *rs, int frame_index, float *noise_rgba,
                        const float *adjoint_rgb, const ymrt_adjoint_targets *targets, void *stream, int timing,
                        ymrt_frame_stats *stats)
{
    YM_REQUIRE(sc != nullptr && rs != nullptr && noise_rgba != nullptr, "NULL argument");
    YM_REQUIRE(adjoint_rgb != nullptr && targets != nullptr, "NULL adjoint buffer or targets");
    YM_REQUIRE(frame_index >= 1, "frame_index starts at 1");
    try {
        YM_HIP(hipSetDevice(sc->ctx->device));
        AdjParams adj;
        int rc = prepare_adjoint(sc, targets, adjoint_rgb, (hipStream_t)stream, &adj);
        if (rc) return rc;
        rc = render_frame(sc, rs, frame_index, noise_rgba, nullptr, &adj, (hipStream_t)stream, timing, stats);
        if (rc) return rc;
    } catch (const std::exception &e) {
        set_error(std::string("render failed: ") + e.what());
        return YMRT_ERR_INTERNAL;
    }
    return check_stack_flag(sc);
}

int ymrt_render_frame_aov(ymrt_scene *sc, const ymrt_render_settings *rs, int frame_index, float *noise_rgba,
                          uint32_t *instance_ids, void *stream, int timing, ymrt_frame_stats *stats)
{
    YM_REQUIRE(sc != nullptr && rs != nullptr && noise_rgba != nullptr, "NULL argument");
    YM_REQUIRE(frame_index >= 1, "frame_index starts at 1");
    try {
        int rc = render_frame(sc, rs, frame_index, noise_rgba, instance_ids, nullptr, (hipStream_t)stream,
                              timing, stats);
        if (rc) return rc;
    } catch (const std::exception &e) {
        set_error(std::string("render failed: ") + e.what());
        return YMRT_ERR_INTERNAL;
    }
    return check_stack_flag(sc);
}

int ymrt_accumulate(const float *noise, float *accum, uint32_t w, uint32_t h, int frame, int max_frames,
                    void *stream)
{
    YM_REQUIRE(noise && accum, "NULL buffer");
    YM_REQUIRE(frame >= 1, "frame_index starts at 1");
    if (max_frames > 0 && frame - 1 >= max_frames) return YMRT_OK;
    return accumulate(noise, accum, w, h, frame, (hipStream_t)stream);
}

int ymrt_tonemap(float *rgba, uint32_t w, uint32_t h, float gamma, float exposure, void *stream)
{
    YM_REQUIRE(rgba, "NULL buffer");
    YM_REQUIRE(gamma > 0.0f, "gamma must be positive");
    return tonemap(rgba, w, h, gamma, exposure, (hipStream_t)stream);
}

int ymrt_halton_values(ymrt_scene *sc, const uint64_t *indices, uint32_t n, uint32_t dim0, uint32_t dims,
                       float *out, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_REQUIRE(n == 0 || (indices && out), "NULL buffer");
    int rc = launch_halton_values(sc, indices, n, dim0, dims, out, (hipStream_t)stream);
    if (rc) return rc;
    YM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return YMRT_OK;
}

int ymrt_render_batch(ymrt_context *ctx, const ymrt_batch_item *items, uint32_t n, uint32_t max_slots,
                      void *stream)
{
    try {
        return render_batch(ctx, items, n, max_slots, (hipStream_t)stream);
    } catch (const std::exception &e) {
        set_error(std::string("batch failed: ") + e.what());
        return YMRT_ERR_INTERNAL;
    }
}

int ymrt_autotune(ymrt_scene *sc, const ymrt_render_settings *rs, int frame_index, void *stream,
                  ymrt_tune_result *result)
{
    const int rc = autotune(sc, rs, frame_index, (hipStream_t)stream, result);
    return rc ? rc : check_stack_flag(sc);
}

int ymrt_scene_update_materials(ymrt_scene *sc, uint32_t first, uint32_t count,
                                const ymrt_material_desc *materials, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    return update_materials(sc, first, count, materials, (hipStream_t)stream);
}

int ymrt_scene_update_texture(ymrt_scene *sc, uint32_t texture, const float *texels, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_HIP(hipSetDevice(sc->ctx->device));
    return update_texture(sc, texture, texels, (hipStream_t)stream);
}

int ymrt_scene_update_grid(ymrt_scene *sc, uint32_t grid, const float *density, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_HIP(hipSetDevice(sc->ctx->device));
    return update_grid(sc, grid, density, (hipStream_t)stream);
}

int ymrt_texture_eval(ymrt_scene *sc, uint32_t texture, const float *coords, uint32_t n, float *out, void *stream)
{
    YM_REQUIRE(sc != nullptr, "scene is NULL");
    YM_REQUIRE(n == 0 || (coords && out), "NULL buffer");
    int rc = launch_texture_eval(sc->dev, texture, coords, n, out, (hipStream_t)stream);
    if (rc) return rc;
    YM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return YMRT_OK;
}

int ymrt_halton_enumerate(uint32_t width, uint32_t height, const uint32_t *px, const uint32_t *py,
                          const uint32_t *sample, uint32_t n, uint64_t *out, void *stream)
{
    YM_REQUIRE(n == 0 || (px && py && sample && out), "NULL buffer");
    int rc = launch_halton_enumerate(width, height, px, py, sample, n, out, (hipStream_t)stream);
    if (rc) return rc;
    YM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return YMRT_OK;
}

}  // extern "C"

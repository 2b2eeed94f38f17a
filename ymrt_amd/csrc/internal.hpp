// internal.hpp -- library-private types behind the C ABI (include/ymrt.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ymrt.h"
#include "dev_scene.hpp"

namespace ym {

void set_error(const std::string &msg);

#define YM_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            ::ym::set_error(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" __FILE__ ":" +    \
                            std::to_string(__LINE__) + ")");                                           \
            return e_ == hipErrorOutOfMemory ? YMRT_ERR_OUT_OF_MEMORY : YMRT_ERR_DEVICE;               \
        }                                                                                              \
    } while (0)

#define YM_REQUIRE(cond, msg)                                                                          \
    do {                                                                                               \
        if (!(cond)) {                                                                                 \
            ::ym::set_error(msg);                                                                      \
            return YMRT_ERR_INVALID_ARGUMENT;                                                          \
        }                                                                                              \
    } while (0)

struct RenderBuffers;  // buffers.hpp
struct AdjParams;      // pt.hpp
struct BatchMemory;    // batch.cpp
void free_batch_memory(BatchMemory *m);

}  // namespace ym

struct ymrt_context {
    int device = 0;
    ym::BatchMemory *batch = nullptr;  // ymrt_render_batch's buffers
    std::mutex batch_mu;
    ~ymrt_context() { ym::free_batch_memory(batch); }
};

struct ymrt_scene {
    ymrt_context *ctx = nullptr;
    ym::DevScene dev{};
    std::vector<void *> allocations;
    ymrt_scene_info info{};
    uint32_t max_stack = 0;
    bool needs_nearby = false;    // boundaries, dielectrics or ior priorities present
    ym::RenderBuffers *render = nullptr;
    uint32_t *d_error = nullptr;  // device flag: traversal stack overflow
    // host state that ymrt_scene_update_materials re-derives flags from
    std::vector<ym::DevMaterial> materials;
    std::vector<ym::DevTexture> textures;  // texture records (float texel updates)
    std::vector<ym::DevVolume> volumes;    // volume and grid records (adjoint targets)
    std::vector<ym::DevGrid> grids;
    std::vector<ym::DevLightRec> light_recs;  // shape-light sampling records, whose emission
    std::vector<uint32_t> light_material;     // follows updates of these materials
    int32_t *d_adj_slots = nullptr;        // adjoint: 2 x material_count target slots
    uint32_t texture_count = 0;
    bool has_boundaries = false;
    bool procedural_textures = false;
};

namespace ym {

// kernels_trace.hip
int launch_trace(const ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits, uint32_t *occluded,
                 uint32_t *visits, int any, hipStream_t stream);
int launch_halton_values(const ymrt_scene *sc, const uint64_t *idx, uint32_t n, uint32_t dim0, uint32_t dims,
                         float *out, hipStream_t stream);
int launch_halton_enumerate(uint32_t w, uint32_t h, const uint32_t *px, const uint32_t *py, const uint32_t *s,
                            uint32_t n, uint64_t *out, hipStream_t stream);
int upload_sampler_tables(ymrt_scene *sc);

// kernels_pt.hip
int launch_texture_eval(const DevScene &sc, uint32_t tex, const float *coords, uint32_t n, float *out,
                        hipStream_t stream);

// render.cpp
void destroy_render_buffers(ymrt_scene *sc);
void free_render_buffers(RenderBuffers *rb);
// One frame with the path-state buffers in `buffers` (allocated or grown as
// the frame needs; the scene's own for the single-frame entry points).
// adj: the adjoint's targets (a replay pass follows each chunk), or null.
int render_frame_with(ymrt_scene *sc, RenderBuffers *&buffers, const ymrt_render_settings *rs, int frame,
                      float *noise, uint32_t *instance_ids, const AdjParams *adj, hipStream_t stream, int timing,
                      ymrt_frame_stats *stats);
// adjoint.cpp: checks the targets of ymrt_render_adjoint against the scene and
// fills the replay's parameters (uploading the per-material target slots)
int prepare_adjoint(ymrt_scene *sc, const ymrt_adjoint_targets *tg, const float *adj_img, hipStream_t stream,
                    AdjParams *out);

// abi.cpp: the traversal-overflow flag of a scene, cleared before a call's
// kernels and read once its stream has drained
int clear_stack_flag(ymrt_scene *sc, hipStream_t stream);
int check_stack_flag(ymrt_scene *sc);

}  // namespace ym

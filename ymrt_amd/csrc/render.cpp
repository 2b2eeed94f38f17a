// render.cpp -- frame driver (filled in with the wavefront path tracer).
#include "internal.hpp"

namespace ym {

struct RenderBuffers {};

void destroy_render_buffers(ymrt_scene *sc)
{
    delete sc->render;
    sc->render = nullptr;
}

int render_frame(ymrt_scene *, const ymrt_render_settings *, int, float *, hipStream_t, int, ymrt_frame_stats *)
{
    set_error("path tracing not built yet");
    return YMRT_ERR_UNSUPPORTED;
}
int accumulate(const float *, float *, uint32_t, uint32_t, int, hipStream_t)
{
    set_error("not built yet");
    return YMRT_ERR_UNSUPPORTED;
}
int tonemap(float *, uint32_t, uint32_t, float, float, hipStream_t)
{
    set_error("not built yet");
    return YMRT_ERR_UNSUPPORTED;
}

}  // namespace ym

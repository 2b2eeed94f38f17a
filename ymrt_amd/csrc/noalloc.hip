// noalloc.hip -- ymrt_render_frame_noalloc: a frame that allocates no device
// memory, for callers that cannot call hipMalloc (code under stream capture,
// tight serving loops, harnesses that forbid allocation).
//
// The path state lives in `scratch`, a fixed device array of
// kNoAllocScratchBytes that every GPU running the library holds from module
// load until the process ends.  Host code never passes the array's name to
// a kernel or a copy: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device when a context is created
// (ymrt_context_create) and kept in g_scratch[device].  All calls on one GPU
// share that GPU's array, so calls that overlap on one GPU race; the frame
// clears the pass counters and statistics it reads before its kernels run.
#include <mutex>

#include "internal.hpp"
#include "pt.hpp"

namespace ym {

// 64 MiB: 65536 paths of path state (about 656 B each) and the traversal
// spill rows of 1024 waves (48 entries x 4 B per lane).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr uint32_t kNoAllocPaths = 1u << 16;
constexpr uint32_t kNoAllocSpillThreads = 1024u * 64u;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[kMaxDevices] = {};

size_t fixed_buffer_bytes(uint32_t cap, uint32_t spill_threads);  // render.cpp
int render_frame_fixed(ymrt_scene *sc, void *memory, size_t bytes, uint32_t cap, uint32_t spill_threads,
                       const ymrt_render_settings *rs, int frame, float *noise, hipStream_t stream,
                       ymrt_frame_stats *stats);  // render.cpp

// Called with `device` current (ymrt_context_create).
int noalloc_setup(int device)
{
    YM_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the fixed-memory table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return YMRT_OK;
    void *ptr = nullptr;
    YM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    YM_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for the fixed device memory");
    g_scratch[device] = ptr;
    return YMRT_OK;
}

int render_frame_noalloc(ymrt_scene *sc, const ymrt_render_settings *rs, int frame, float *noise,
                         hipStream_t stream, ymrt_frame_stats *stats)
{
    const int dev = sc->ctx->device;
    YM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the fixed-memory table");
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ptr = g_scratch[dev];
    }
    YM_REQUIRE(ptr != nullptr, "fixed device memory not set up on this device (create the context first)");
    static_assert(kNoAllocPaths % 256u == 0, "whole tiles of paths");
    YM_REQUIRE(fixed_buffer_bytes(kNoAllocPaths, kNoAllocSpillThreads) <= kNoAllocScratchBytes,
               "fixed path state exceeds the fixed device memory");
    return render_frame_fixed(sc, ptr, kNoAllocScratchBytes, kNoAllocPaths, kNoAllocSpillThreads, rs, frame, noise,
                              stream, stats);
}

}  // namespace ym

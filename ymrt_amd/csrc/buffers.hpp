// buffers.hpp -- path-state buffers of the frame drivers (render.cpp, batch.cpp).
#pragma once

#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "pt.hpp"

namespace ym {

// Pass counters per chunk slot.
constexpr size_t kCountersPerSlot = (size_t)(kMaxPasses + 2) * kCtrStride;

struct RenderBuffers {
    uint32_t cap = 0;
    // the adjoint's replay pass: its own pass counters, statistics and slot
    uint32_t *adj_counters = nullptr;
    unsigned long long *adj_stats = nullptr;
    FrameSlot *adj_slot = nullptr;
    PathState ps{};
    std::vector<void *> allocs;
    uint32_t *counters = nullptr;
    unsigned long long *stats = nullptr;
    FrameSlot *slots = nullptr;  // device table of the chunks a launch serves
    std::vector<hipEvent_t> events;
    size_t event_next = 0;

    ~RenderBuffers()
    {
        for (void *p : allocs) (void)hipFree(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

template <typename T>
inline int alloc(RenderBuffers *rb, T **p, size_t count)
{
    void *m = nullptr;
    YM_HIP(hipMalloc(&m, sizeof(T) * std::max<size_t>(count, 1)));
    rb->allocs.push_back(m);
    *p = static_cast<T *>(m);
    return YMRT_OK;
}

// Lays out the path state of n paths (and the pass counters and statistics)
// through get(pointer, count), which allocates or carves each array.  A batch
// shares the path state among up to `slots` frame chunks: pass counters and
// compaction counts per slot, statistics per frame (stat_sets).
template <typename Get>
inline int layout_buffers(RenderBuffers &rb, size_t n, uint32_t spill_threads, Get &&get, uint32_t slots = 1,
                          uint32_t stat_sets = 1)
{
    PathState &ps = rb.ps;
    int rc = 0;
    rc |= get(ps.ray_o, n);
    rc |= get(ps.ray_d, n);
    rc |= get(ps.last, n);
    rc |= get(ps.beta, n);
    rc |= get(ps.L, n);
    rc |= get(ps.smp, n);
    rc |= get(ps.info, n);
    rc |= get(ps.xhdr, n);
    rc |= get(ps.xrec, n * kMaxBoundaryRecord);
    rc |= get(ps.hit_a, n);
    rc |= get(ps.hit_b, n);
    rc |= get(ps.nb, n * 4);
    rc |= get(ps.sh_o, n * kShadowSlots);
    rc |= get(ps.sh_d, n * kShadowSlots);
    rc |= get(ps.sh_f, n * kShadowSlots);
    rc |= get(ps.sh_li, n * kShadowSlots);
    rc |= get(ps.sh_tr, n * kShadowSlots);
    rc |= get(ps.sh_c, n);
    rc |= get(ps.sh_flags, n);
    rc |= get(ps.out, n);
    rc |= get(ps.queue[0], n);
    rc |= get(ps.queue[1], n);
    rc |= get(ps.q_hit, n);
    rc |= get(ps.q_hit_vol, n);
    rc |= get(ps.q_hit_diffuse, n);
    rc |= get(ps.q_hit_metal, n);
    rc |= get(ps.q_medium, n);
    rc |= get(ps.flag, n);
    rc |= get(ps.ccount, 4 * (n / kCompactSpan + 1) * slots);
    rc |= get(ps.q_shadow, n);
    rc |= get(ps.q_transmit, n);
    rc |= get(ps.spill, (size_t)(kMaxStackDepth - kStackWindow) * spill_threads);
    rc |= get(ps.first_hit, n);
    rc |= get(rb.counters, kCountersPerSlot * slots);
    rc |= get(rb.stats, (size_t)kStatCount * stat_sets);
    rc |= get(rb.slots, (size_t)slots);
    ps.spill_threads = spill_threads;
    return rc;
}

}  // namespace ym

// pt.hpp -- wavefront path-tracing state shared by render.cpp and kernels_pt.hip.
#pragma once

#include "../../include/ymrt.h"
#include "dev_sampler.hpp"

namespace ym {

// Per-pass device counters: counters[pass * kCtrStride + k]
constexpr uint32_t kCtrActive = 0;       // extension rays queued for this pass
constexpr uint32_t kCtrHit = 1;          // surface hits queued for shading
constexpr uint32_t kCtrShadow = 2;       // paths with shadow rays
constexpr uint32_t kCtrHitVol = 3;       // volume scattering vertices queued for shading
constexpr uint32_t kCtrWorkRaygen = 4;   // work-fetch cursors
constexpr uint32_t kCtrWorkExtend = 5;
constexpr uint32_t kCtrWorkShade = 6;
constexpr uint32_t kCtrWorkShadeVol = 7;
constexpr uint32_t kCtrWorkShadow = 8;
constexpr uint32_t kCtrWorkFlush = 9;
constexpr uint32_t kCtrMedium = 10;        // paths inside media queued for free flight
constexpr uint32_t kCtrWorkMedium = 11;
constexpr uint32_t kCtrTransmit = 12;      // paths whose unoccluded shadow rays cross media
constexpr uint32_t kCtrWorkTransmit = 13;
constexpr uint32_t kCtrHitDiffuse = 14;    // surface hits on diffuse-only materials
constexpr uint32_t kCtrWorkShadeDiffuse = 15;
constexpr uint32_t kCtrHitMetal = 16;      // surface hits on metal-only materials
constexpr uint32_t kCtrWorkShadeMetal = 17;
constexpr uint32_t kCtrStride = 24;

// Compaction: blocks of kCompactThreads threads, kCompactSpan queue entries
// each.
constexpr uint32_t kCompactThreads = 256;
constexpr uint32_t kCompactSpan = 4096;

// Stage flags: flag[slot] = (pass << kFlagBits) | kind names the stage a
// path goes to next in that pass; the compaction kernels gather each kind
// into its queue in slot order.  Flags of earlier passes are stale and ignored;
// raygen clears every slot of a chunk.
constexpr uint32_t kFlagBits = 3;
constexpr uint32_t kFlagHit = 1;      // surface shading
constexpr uint32_t kFlagMedium = 2;   // free flight
constexpr uint32_t kFlagVolume = 3;   // volume-vertex shading
constexpr uint32_t kFlagDiffuse = 4;  // surface shading, diffuse-only material
constexpr uint32_t kFlagTransmit = 5; // transmittance of unoccluded shadow rays
constexpr uint32_t kFlagMetal = 6;    // surface shading, metal-only material
constexpr uint32_t kMaxPasses = 4 * kMaxRayDepth;

// Path-tracing traversals keep the top kStackWindow stack entries per lane in
// LDS and spill deeper ones to a global area of (kMaxStackDepth - kStackWindow)
// entries for each of kSpillThreads threads (the largest persistent grid:
// 5 waves on each SIMD of 256 CUs, the most the shadow kernel's registers
// allow; 63 MB).
constexpr int kStackWindow = 16;
constexpr int kMaxStackDepth = 64;
constexpr uint32_t kSpillThreads = 256u * 4u * 5u * 64u;
// Most waves (64 lanes) of a persistent pass kernel: one spill row per lane.
constexpr uint32_t kMaxPersistentWaves = kSpillThreads / 64u;

// Frame statistics (device, 64-bit)
constexpr uint32_t kStatCamera = 0;
constexpr uint32_t kStatExtension = 1;
constexpr uint32_t kStatShadow = 2;
constexpr uint32_t kStatTransmittance = 3;
constexpr uint32_t kStatExtendNodes = 4;   // wide nodes visited by closest-hit traversals
constexpr uint32_t kStatExtendTris = 5;    // triangles tested by closest-hit traversals
constexpr uint32_t kStatShadowNodes = 6;   // same for any-hit shadow traversals
constexpr uint32_t kStatShadowTris = 7;
constexpr uint32_t kStatTransmitNodes = 8; // boundary walks of shadow rays in media
constexpr uint32_t kStatTransmitTris = 9;
constexpr uint32_t kStatCount = 12;

// Structure-of-arrays path state for `cap` paths in flight (indexed by slot).
struct PathState {
    float4 *ray_o, *ray_d;
    uint2 *last;
    float4 *beta, *L;
    uint4 *smp;
    uint32_t *info;
    uint32_t *xhdr;
    uint4 *xrec;        // [kMaxBoundaryRecord][cap]
    uint4 *hit_a;       // t bits, instance, triangle, flags
    float4 *hit_b;      // barycentrics or volume position
    uint2 *nb;          // [4][cap] nearby (instance | back << 31, t bits)
    float4 *sh_o, *sh_d, *sh_f, *sh_li;  // [kShadowSlots][cap]
    float4 *sh_tr;      // [kShadowSlots][cap] transmittance of shadow rays in media
    float4 *sh_c;       // throughput at the vertex, light-select pdf
    uint32_t *sh_flags;
    float4 *out;        // final radiance per path
    uint32_t *queue[2];
    uint32_t *q_hit, *q_hit_vol, *q_hit_diffuse, *q_hit_metal, *q_medium, *q_shadow, *q_transmit;
    uint32_t *spill;    // traversal stack spill [kMaxStackDepth - kStackWindow][spill_threads]
    uint32_t spill_threads;  // spill rows: the most lanes a persistent pass kernel may launch
    uint32_t *flag;       // stage flag per slot (kFlag*)
    uint32_t *ccount;     // compaction: matches per block, [4][cap / kCompactSpan + 1]
    uint32_t *first_hit;  // instance hit by the camera ray (written when an id image is requested)
    uint2 *adj_g;         // adjoint replay: (roughness entry or kEmpty, d ln f / d r of the
                          // continuation direction) at the path's current surface vertex
};

// Shadow rays per path vertex: slots 0-1 shape light (light sample, BSDF or
// phase sample), slots 2-3 distant light.
constexpr int kShadowSlots = 4;

// Targets of the path-replay adjoint (ymrt_render_adjoint) in the replay
// pass: gradient events add to these device arrays.
constexpr uint32_t kAdjLds = 256;  // albedo rows x 3 + roughness entries summed in LDS first

struct AdjParams {
    const float *img;          // dLoss/dImage, width * height * 3
    const int32_t *alb_slot;   // per material: row of g_alb, or -1
    const int32_t *rough_slot; // per material: entry of g_rough, or -1
    float *g_alb, *g_rough, *g_tex, *g_vox;
    uint32_t n_alb, n_rough;
    uint32_t tex;              // target image texture, or kEmpty
    uint32_t vol;              // target volume (gray sigma_t, voxel grid), or kEmpty
    uint32_t lds;              // 1: albedo and roughness sums fit kAdjLds floats
};

struct PtParams {
    ymrt_render_settings rs;
    HaltonEnum henum;
    int frame;
    uint32_t tiles_x;
    uint64_t chunk_base;     // first (virtual pixel * spp + sample) of the chunk
    uint32_t chunk_paths;
    uint32_t cap;
    uint32_t *counters;
    unsigned long long *stats;
    uint32_t *error;
    uint32_t *aov;           // per-pixel instance ids of the camera rays, or null
    AdjParams adj;           // the replay pass of ymrt_render_adjoint (zero otherwise)
    uint32_t ccount_stride;  // entries per queue in ccount
    uint32_t media;          // volume scattering on and the scene has volumes
    float4 *noise;           // the frame's noise image (multi-frame resolve)
};

// Scene features a multi-frame launch compiles in: the union over its slots.
struct MultiVariant {
    bool nearby, media, sphere, graph;
    int het;  // 0 homogeneous media, 1 voxel grids, 2 texture densities
    bool adj = false;  // the adjoint's replay pass: kernels that emit gradient events
};

// Most frame chunks one multi-frame launch serves.
constexpr uint32_t kMaxSlots = 64;

// One frame chunk of a multi-frame launch: scene, parameters and the slot's
// part of the path state (device-resident table, read with scalar loads).
struct FrameSlot {
    DevScene sc;
    PtParams prm;
    PathState ps;
};

}  // namespace ym

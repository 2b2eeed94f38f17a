/*
 * ymrt.h -- C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * Plain C: POD structs, opaque handles, int status codes, device pointers and
 * a hipStream_t passed as void*.  No C++ or GPU-library types cross this
 * boundary, and no exception leaves it.  On failure every entry point returns a
 * non-zero YMRT_ERR_* code and ymrt_last_error() holds a message for the
 * calling thread.
 *
 * Reference interfaces replaced (paths relative to the YMRT source tree):
 *   - Scene upload: SceneModule::GetHostSceneDataPointer
 *     (src/core/SceneModule.cpp:1434-1463) and the host steps of
 *     GuiModule::UpdateScene (src/core/GuiModule.cpp:116-271): top BVH build,
 *     shape-light CDF (SceneModule::LoadShapeLight, SceneModule.cpp:1203-1304),
 *     Halton permutation table (SceneModule.cpp:1412-1432), camera medium
 *     (Camera::InitCameraRayTransfer, src/Camera.cpp:15-78), UPLOAD_TO_GPU.
 *   - Ray queries: TraceRay (src/ShadingUtilities.cuh:310-431) and
 *     BVHTraverseShadow (src/ShadingUtilities.cuh:563-635).
 *   - Frame: the extern "C" stage launchers RayGenerationPathTracing
 *     (src/Shading.cu:1727), RayTracePathTracing (1903), RayShadingPathTracing
 *     (2304) driven by RenderModule::PathTracingExecuteTask
 *     (src/core/RenderModule.cpp:494-653); AccumulateImagePathTracing (2375);
 *     PostProcessingPathTracing (2502).
 */
#ifndef YMRT_H
#define YMRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YMRT_ABI_VERSION 8

/* status codes */
#define YMRT_OK 0
#define YMRT_ERR_INVALID_ARGUMENT 1
#define YMRT_ERR_DEVICE 2
#define YMRT_ERR_OUT_OF_MEMORY 3
#define YMRT_ERR_UNSUPPORTED 4
#define YMRT_ERR_INTERNAL 5

#define YMRT_INVALID_INDEX 0xFFFFFFFFu

/* ---------------------------------------------------------------- scene data
 * These descriptors carry the same data as YMRT's host arrays: a mesh is the
 * TriangleMesh blob of src/GeometryDefines.h:197-273 (per-corner position,
 * normal and texcoord indices), an instance is PrimitiveInstance
 * (src/PrimtiveInstance.h:9-25) with its transform pair (src/SceneDefines.h:84-86),
 * a material is Material/DefaultMtl/LightMtl (src/Material.h:21-359), a volume
 * is Volume (src/Volume.h:9-18).  Matrices are column-major 4x4 like glm::mat4.
 */

typedef struct ymrt_mesh_desc {
    uint32_t triangle_count;
    uint32_t position_count;         /* number of float3 positions           */
    const uint32_t *position_indices; /* 3 * triangle_count                   */
    const float *positions;          /* 3 * position_count                    */
    uint32_t normal_count;
    const uint32_t *normal_indices;  /* 3 * triangle_count, or NULL with normals NULL */
    const float *normals;            /* 3 * normal_count                      */
    uint32_t texcoord_count;
    const uint32_t *texcoord_indices; /* 3 * triangle_count, or NULL          */
    const float *texcoords;          /* 2 * texcoord_count                    */
} ymrt_mesh_desc;

/* Analytic sphere centred at the object-space origin (Sphere,
 * src/GeometryDefines.h; intersected by IntersectSphereObject,
 * src/ShadingUtilities.cuh:219-275). */
typedef struct ymrt_sphere_desc {
    float radius;
} ymrt_sphere_desc;

#define YMRT_GEOMETRY_MESH 0u
#define YMRT_GEOMETRY_SPHERE 1u

typedef struct ymrt_instance_desc {
    uint32_t geometry;               /* index into meshes or spheres          */
    uint32_t material;               /* index into materials                  */
    int32_t inner_volume;            /* index into volumes or -1              */
    int32_t treat_as_boundary;       /* 1: medium boundary, skipped by shading */
    uint32_t geometry_type;          /* YMRT_GEOMETRY_MESH or _SPHERE         */
    uint32_t reserved;
    float object_to_world[16];
    float world_to_object[16];
} ymrt_instance_desc;

#define YMRT_MATERIAL_DEFAULT 0u
#define YMRT_MATERIAL_LIGHT 1u

typedef struct ymrt_material_desc {
    uint32_t type;
    /* DefaultMtl (roughness values are squared into GGX alpha at shading) */
    float diffuse_albedo[3];   uint32_t diffuse_albedo_tex;
    float specular_albedo[3];  uint32_t specular_albedo_tex;
    float roughness_x;         uint32_t roughness_x_tex;
    float roughness_y;         uint32_t roughness_y_tex;
    float specular_weight;     uint32_t specular_weight_tex;
    float metalness;           uint32_t metalness_tex;
    float transmission_weight; uint32_t transmission_weight_tex;
    float ior_n;               uint32_t ior_priority;
    float coat_albedo[3];      uint32_t coat_albedo_tex;
    float coat_weight;         uint32_t coat_weight_tex;
    float coat_thickness;      uint32_t coat_thickness_tex;
    float coat_roughness_x;    uint32_t coat_roughness_x_tex;
    float coat_roughness_y;    uint32_t coat_roughness_y_tex;
    float coat_ior;
    uint32_t normal_mapping_tex;     /* tangent-space normal map (rgb*2-1)    */
    uint32_t bump_mapping_tex;       /* height map (x)                        */
    /* LightMtl */
    float light_color[3];
    float intensity;
} ymrt_material_desc;

#define YMRT_WRAP_REPEAT 0
#define YMRT_WRAP_CLAMP 1
#define YMRT_COLOR_SRGB 0
#define YMRT_COLOR_LINEAR 1

/* Texture node (Texture, src/Texture.h:16-720).  IMAGE: 8-bit texels decoded
 * like LoadPng8 (src/FileIO.cpp:17-45): value/255, then sRGB->linear when
 * color_space == YMRT_COLOR_SRGB; row 0 is the bottom row; channels 1..4.
 * CONSTANT_*: value.  The procedural nodes (TextureEval.cuh:238-404) blend
 * two inputs child[0] ("black") and child[1] ("white"), each another texture
 * node or YMRT_INVALID_INDEX for black; a graph may nest 8 deep. */
#define YMRT_TEXTURE_IMAGE 0u
#define YMRT_TEXTURE_CONSTANT_FLOAT 1u
#define YMRT_TEXTURE_CONSTANT_RGB 2u
#define YMRT_TEXTURE_CHECKERBOARD 3u
#define YMRT_TEXTURE_NOISE 4u
#define YMRT_TEXTURE_FBM 5u
#define YMRT_TEXTURE_TURBULENCE 6u
#define YMRT_TEXTURE_MARBLE 7u
#define YMRT_TEXTURE_WOOD 8u
#define YMRT_TEXTURE_POLKA_DOT 9u
#define YMRT_TEXTURE_WAVE 10u

typedef struct ymrt_texture_desc {
    uint32_t width, height, channels;
    const uint8_t *texels;           /* width * height * channels             */
    int32_t wrap_mode;
    int32_t color_space;
    float scale_u, scale_v, offset_u, offset_v;
    uint32_t type;                   /* YMRT_TEXTURE_*                        */
    float value[3];                  /* CONSTANT_FLOAT uses value[0]         */
    uint32_t child[2];               /* procedural inputs: black, white       */
    float frequency, lacunarity, gain, amplitude, offset, variation, radius;
    int32_t layer_count;
    int32_t normalized;              /* NOISE: remap [-1,1] to [0,1]          */
    int32_t turbulence_axes;         /* MARBLE: bit 0 x, bit 1 y, bit 2 z     */
    float frequency1, lacunarity1, gain1;   /* WAVE: second fBm (the first uses */
    int32_t layer_count1;                    /* frequency, lacunarity, gain)     */
    /* IMAGE with float texels: width * height * channels linear values (the
     * FP32 image tiles the reference decodes its files into,
     * src/TextureManager.h:19-50); when set, texels and color_space are
     * unused.  These are the texels ymrt_render_adjoint differentiates. */
    const float *texels_f32;
} ymrt_texture_desc;

/* Density grid for heterogeneous media (no counterpart in YMRT, whose
 * heterogeneous density comes only from its texture graph; see
 * ymrt_volume_desc.density_texture).  Voxel-constant density over the unit
 * cube [0,1]^3 of the volume's object space. */
typedef struct ymrt_grid_desc {
    uint32_t nx, ny, nz;
    const float *density;            /* nx*ny*nz, x fastest                   */
} ymrt_grid_desc;

/* Volume (src/Volume.h:9-18).  The density is homogeneous (1), a voxel grid
 * or, as YMRT's density_texture_idx, the x channel of a texture graph
 * evaluated at (st = 0, world point, object point), clamped at 0, under the
 * majorant Texture::GetUpperBound (src/Texture.h:1181-1251) -- the tracker
 * of SampleTrHeterogeneous / EvalTrHeterogeneous (src/VolumeEval.cuh:80-418).
 * At most one of density_grid and density_texture is >= 0. */
typedef struct ymrt_volume_desc {
    float sigma_t[3];
    float albedo[3];
    float g;                         /* Henyey-Greenstein asymmetry           */
    int32_t density_grid;            /* -1: no grid                           */
    float world_to_object[16];       /* the volume's object space (the grid's
                                        unit cube; the texture's object point) */
    int32_t density_texture;         /* -1: no density texture                */
} ymrt_volume_desc;

/* DistantLight (src/Light.h:324-386): radiance light_color * intensity from
 * the directions within theta_max degrees of -y of the light's frame. */
typedef struct ymrt_distant_light_desc {
    float light_color[3];
    float intensity;
    float theta_max;                 /* cone half-angle, degrees              */
    float object_to_world[16];       /* the light's frame                     */
    float world_to_object[16];
} ymrt_distant_light_desc;

typedef struct ymrt_camera_desc {
    float position[3];
    float w[3];                      /* backward axis, Camera::SetDir         */
    float fov;                       /* vertical field of view, radians       */
    float aspect_ratio;
} ymrt_camera_desc;

typedef struct ymrt_scene_desc {
    uint32_t mesh_count;       const ymrt_mesh_desc *meshes;
    uint32_t sphere_count;     const ymrt_sphere_desc *spheres;
    uint32_t instance_count;   const ymrt_instance_desc *instances;
    uint32_t material_count;   const ymrt_material_desc *materials;
    uint32_t texture_count;    const ymrt_texture_desc *textures;
    uint32_t grid_count;       const ymrt_grid_desc *grids;
    uint32_t volume_count;     const ymrt_volume_desc *volumes;
    uint32_t distant_light_count; const ymrt_distant_light_desc *distant_lights;
    ymrt_camera_desc camera;
    uint32_t halton_seed;      /* seed of the Halton digit permutations     */
} ymrt_scene_desc;

/* ---------------------------------------------------------------- ray queries */

/* A query ray: hits count for t in [tmin, tmax] along d as given.  The
 * reference's rays have no tmin (TraceRay starts at TMIN = 1e-7); pass 1e-7
 * for its answers. */
typedef struct ymrt_ray {          /* 32 B */
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} ymrt_ray;

typedef struct ymrt_hit {          /* 16 B */
    float t;                         /* hit distance along d (d as given)     */
    uint32_t triangle;               /* index in the mesh's triangle list, or YMRT_INVALID_INDEX */
    uint32_t instance;               /* index in the instance list            */
    uint32_t back;                   /* 1 when d points along the face normal */
} ymrt_hit;

/* ---------------------------------------------------------------- rendering */

#define YMRT_SAMPLER_PCG 0
#define YMRT_SAMPLER_HALTON 1

/* Mirrors RenderSetting (src/SceneDefines.h:17-40). */
typedef struct ymrt_render_settings {
    uint32_t width, height;
    int32_t spp;
    int32_t ray_depth;
    int32_t sampler_type;
    int32_t enable_shape_light;
    int32_t enable_distant_light;
    int32_t enable_volume_scattering;
    int32_t enable_env_light;
    int32_t enable_russian_roulette;
    int32_t enable_indirect_clamp;
    float indirect_limit;
    float gamma;
    float exposure;
    uint32_t path_capacity;          /* paths in flight per chunk; 0: up to 64 M,
                                        within a quarter of free device memory
                                        (about 720 B per path)                */
    uint32_t persistent_waves;       /* waves of each persistent pass kernel,
                                        at most 5120 (the traversal spill
                                        area); 0: one per 256 paths of the
                                        chunk, within [1024, 5120].  Each
                                        kernel launches at most the waves the
                                        GPU holds of it at once.  Like
                                        path_capacity it changes no pixel.    */
} ymrt_render_settings;

typedef struct ymrt_frame_stats {
    uint64_t camera_rays;            /* extension segments with depth 1       */
    uint64_t extension_rays;         /* all closest-hit segments              */
    uint64_t shadow_rays;            /* any-hit segments                      */
    uint64_t transmittance_rays;     /* boundary segments of shadow transmittance */
    uint64_t paths;
    uint32_t passes;                 /* bounce passes launched                */
    uint32_t chunks;
    /* Per-stage GPU time measured with HIP events on the render stream (ms,
     * summed over launches) and launch counts; filled when timing != 0. */
    float ms_raygen, ms_extend, ms_shade, ms_shadow, ms_resolve;
    uint32_t launches_extend, launches_shade, launches_shadow;
    float ms_total;
    float ms_medium;                 /* free-flight sampling (scenes with media) */
    uint32_t launches_medium;
    /* Traversal work, always counted: wide nodes visited and triangles
     * tested by closest-hit (extend) and any-hit (shadow) queries. */
    uint64_t extend_nodes, extend_tris, shadow_nodes, shadow_tris;
    float ms_transmit;               /* transmittance of shadow rays in media */
    uint32_t launches_transmit;
    uint64_t transmit_nodes, transmit_tris;   /* boundary walks of those rays */
    float ms_adjoint;                /* the replay pass of ymrt_render_adjoint    */
    uint32_t reserved;
} ymrt_frame_stats;

/* ---------------------------------------------------------------- entry points */

typedef struct ymrt_context ymrt_context;
typedef struct ymrt_scene ymrt_scene;

const char *ymrt_last_error(void);
int ymrt_abi_version(void);

int ymrt_context_create(int device, ymrt_context **out);
int ymrt_context_destroy(ymrt_context *ctx);

/* Builds the two-level wide BVH on the host and uploads every array. */
int ymrt_scene_create(ymrt_context *ctx, const ymrt_scene_desc *desc, ymrt_scene **out);
int ymrt_scene_destroy(ymrt_scene *scene);

/* Scene statistics: bvh node count, max stack depth, bytes resident in HBM. */
typedef struct ymrt_scene_info {
    uint64_t triangle_count;
    uint64_t node_count;             /* wide nodes, bottom + top              */
    uint64_t device_bytes;
    uint32_t max_stack;              /* stack entries a traversal may need    */
    uint32_t shape_light_count;
    float build_ms;
    uint32_t camera_medium_records;
} ymrt_scene_info;
int ymrt_scene_get_info(const ymrt_scene *scene, ymrt_scene_info *info);

/* Replaces materials [first, first + count) of a built scene in HBM
 * (SceneModule::UpdateMaterial, src/core/SceneModule.cpp:654-663, which the
 * editor calls after changing a material).  Synchronous on stream.  As in
 * the reference, the shape-light CDF and the camera's medium records keep
 * the values the scene was built with, while a light material's new colour
 * and intensity are what its lights emit from the next frame on; a material
 * may not change its type (YMRT_ERR_INVALID_ARGUMENT). */
int ymrt_scene_update_materials(ymrt_scene *scene, uint32_t first, uint32_t count,
                                const ymrt_material_desc *materials, void *stream);

/* Releases the path-state buffers the scene's frames allocated (the next
 * frame allocates them again); the scene itself stays.  For callers that
 * render many scenes in turn and keep them resident: RenderModule frees its
 * per-task device memory the same way (src/core/RenderModule.cpp:535-539,
 * 640-653).  Waits for the device. */
int ymrt_scene_release_buffers(ymrt_scene *scene);

/* Replaces the texels of a float image texture (texels_f32 at scene
 * creation) with `texels` (host, width * height * channels linear values),
 * as SceneModule::UpdateTexture does for the editor; synchronous on stream.
 * An optimisation loop over ymrt_render_adjoint's texture gradient writes
 * its updated texels back this way. */
int ymrt_scene_update_texture(ymrt_scene *scene, uint32_t texture, const float *texels, void *stream);

/* Replaces the densities of voxel grid `grid` with `density` (host, nz * ny
 * * nx values, the layout of ymrt_grid_desc) and its majorant with their
 * maximum; synchronous on stream.  No counterpart in the reference, which
 * has no voxel grids; an optimisation loop over ymrt_render_adjoint's
 * density gradient writes its updated grid back this way. */
int ymrt_scene_update_grid(ymrt_scene *scene, uint32_t grid, const float *density, void *stream);

/* Closest hit (TraceRay) in [tmin, tmax] for n device-resident rays. */
int ymrt_trace_closest(ymrt_scene *scene, const ymrt_ray *rays, uint32_t n,
                       ymrt_hit *hits, void *stream);
/* Any hit (BVHTraverseShadow): occluded[i] = 1 when some non-boundary
 * triangle lies in [tmin, tmax]. */
int ymrt_trace_any(ymrt_scene *scene, const ymrt_ray *rays, uint32_t n,
                   uint32_t *occluded, void *stream);
/* Same queries, also counting visited wide nodes and tested triangles per ray
 * (visits[2*i], visits[2*i+1]); used to count the bytes a traversal touches. */
int ymrt_trace_closest_counted(ymrt_scene *scene, const ymrt_ray *rays, uint32_t n,
                               ymrt_hit *hits, uint32_t *visits, void *stream);

/* One progressive frame: every pixel gets spp samples of path tracing;
 * noise_rgba (width*height float4, device) receives the per-frame estimate
 * (radiance in rgb, 1 in alpha), as the reference's noise image does.
 * frame_index starts at 1.  stats may be NULL; timing != 0 records per-stage
 * HIP events (adds a synchronisation at the end of the frame). */
int ymrt_render_frame(ymrt_scene *scene, const ymrt_render_settings *settings,
                      int frame_index, float *noise_rgba, void *stream,
                      int timing, ymrt_frame_stats *stats);

/* ymrt_render_frame without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * path state lives in a fixed 64 MiB device array that the library holds on
 * every GPU it runs on until the process ends, so a frame runs in chunks of at
 * most 65536 paths with at most 1024 waves per pass kernel; the image is the
 * one ymrt_render_frame gives, bit for bit.  Rejected with
 * YMRT_ERR_INVALID_ARGUMENT: path_capacity > 65536, spp > 65536,
 * persistent_waves > 1024.  No per-stage timing (stats gets the counts).
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (other threads or streams) race. */
int ymrt_render_frame_noalloc(ymrt_scene *scene, const ymrt_render_settings *settings,
                              int frame_index, float *noise_rgba, void *stream,
                              ymrt_frame_stats *stats);

/* Autotuner for the two launch settings that change no pixel, path_capacity
 * and persistent_waves: renders trial frames of the caller's settings
 * (frame_index, into an image of its own) and times them with HIP events --
 * capacities {library default, 1/2 and 1/4 of the frame's paths} at the
 * default waves, then waves {4096, 2048, 1024} at the best capacity,
 * each the faster of two frames after one warm-up.  The fastest pair goes to
 * result; the caller copies it into its settings.  No counterpart in the
 * reference, which launches with fixed sizes; it is the autotuner its
 * batched-wavefront plan asks for. */
typedef struct ymrt_tune_result {
    uint32_t path_capacity;          /* 0: the library default was fastest   */
    uint32_t persistent_waves;       /* 0: the default was fastest           */
    float ms;                        /* frame time of that pair              */
    float ms_default;                /* frame time with both at 0            */
    uint32_t trials;                 /* frames rendered, warm-up included    */
} ymrt_tune_result;

int ymrt_autotune(ymrt_scene *scene, const ymrt_render_settings *settings,
                  int frame_index, void *stream, ymrt_tune_result *result);

/* ymrt_render_frame plus an instance-id image: instance_ids (width*height
 * uint32, device, may be NULL) receives the index of the instance that
 * sample 0's camera ray hits first, YMRT_INVALID_INDEX on a miss -- the
 * picking buffer RayTraceEditorView fills (src/Shading.cu:1442-1547), where
 * the id written last among the samples wins. */
int ymrt_render_frame_aov(ymrt_scene *scene, const ymrt_render_settings *settings,
                          int frame_index, float *noise_rgba, uint32_t *instance_ids,
                          void *stream, int timing, ymrt_frame_stats *stats);

/* Parameters ymrt_render_adjoint differentiates, and their gradients
 * (device arrays of floats the call adds to; the caller clears them):
 *  - albedo_materials: Lambert-only DefaultMtl materials (specular,
 *    metalness, transmission and coat weights 0, none of them or the diffuse
 *    albedo textured); grad_albedo[k*3 + c] is d/d diffuse_albedo[c] of
 *    albedo_materials[k];
 *  - roughness_materials: metal-only DefaultMtl materials (metalness 1, no
 *    coat, roughness_x == roughness_y, untextured); grad_roughness[k] is
 *    d/d roughness of roughness_materials[k] (both axes move together);
 *  - texture: an IMAGE texture used only as the diffuse albedo of Lambert
 *    materials (otherwise as above); grad_texture has width * height *
 *    channels entries, d/d the linear texel values (float texels, or the
 *    decoded values of 8-bit ones);
 *  - volume: a volume with a gray sigma_t and a voxel-grid density;
 *    grad_density has nz * ny * nx entries, d/d each voxel's density.
 * YMRT_INVALID_INDEX for texture / volume, or a count of 0, leaves that
 * target out (its gradient pointer may be NULL). */
typedef struct ymrt_adjoint_targets {
    const uint32_t *albedo_materials;       /* host, albedo_count entries    */
    uint32_t albedo_count;
    const uint32_t *roughness_materials;    /* host, roughness_count entries */
    uint32_t roughness_count;
    uint32_t texture;
    uint32_t volume;
    float *grad_albedo;                     /* device, albedo_count * 3      */
    float *grad_roughness;                  /* device, roughness_count       */
    float *grad_texture;                    /* device, texels x channels     */
    float *grad_density;                    /* device, voxels                */
} ymrt_adjoint_targets;

/* Path-replay adjoint of one frame: renders the frame like ymrt_render_frame
 * into noise_rgba, then replays every path with the same samples (each
 * random number is keyed by pixel, sample and vertex, so nothing is stored
 * per vertex) and adds to the targets' gradients the derivative of
 * sum_pixels adjoint_rgb . image (adjoint_rgb: device, width * height * 3)
 * with the frame's own samples, sampling pdfs and MIS weights held fixed:
 * a factor of a vertex or a free flight that scales every later deposit of
 * the path adds adj . (L_path - L_before) * dln(factor), a factor of one
 * shadow ray adds adj . deposit * dln(factor) (GGX roughness: d ln(D G2) /
 * dr; voxel density: the delta tracker's collision probabilities sigma /
 * sigma_m and 1 - sigma / sigma_m, and ratio tracking's 1 - sigma /
 * sigma_m).  A target that is not differentiable as listed above is
 * rejected with YMRT_ERR_INVALID_ARGUMENT naming it; Russian roulette and
 * the indirect clamp must be off.  No counterpart in the reference, whose
 * render loop has no derivatives; it is the adjoint its differentiable
 * rendering workload asks for.  Gradients are summed with float atomics, so
 * their last bits vary between runs. */
int ymrt_render_adjoint(ymrt_scene *scene, const ymrt_render_settings *settings,
                        int frame_index, float *noise_rgba, const float *adjoint_rgb,
                        const ymrt_adjoint_targets *targets, void *stream, int timing,
                        ymrt_frame_stats *stats);

/* One frame of one scene in a batch. */
typedef struct ymrt_batch_item {
    ymrt_scene *scene;
    const ymrt_render_settings *settings;
    int frame_index;
    int timing;                      /* unused: frames share their launches   */
    float *noise_rgba;               /* device, width*height float4           */
    ymrt_frame_stats *stats;         /* may be NULL                           */
} ymrt_batch_item;

/* Renders every item as ymrt_render_frame would, bit for bit, in shared
 * launches: the frames are cut into chunks of whole pixels and packed into
 * rounds of up to max_slots chunks (0 or > 64: 64) and up to 64 M paths (within
 * a quarter of free device memory); one launch of each stage kernel serves
 * every chunk of a round, so small frames fill the GPU together and share one
 * tail of long paths.  Replaces YMRT's render threads, one stream and frame
 * loop per view (src/core/RenderModule.cpp:55-66, 95-143, host loop 580-631).
 * Runs on `stream` and returns when every frame is done; stats get the ray
 * and traversal counts (no per-stage timing: stages are shared).  The
 * round's buffers belong to the context (kept between calls, released by
 * ymrt_context_destroy); calls on one context run one at a time.  On failure
 * the message names the item. */
int ymrt_render_batch(ymrt_context *ctx, const ymrt_batch_item *items, uint32_t n,
                      uint32_t max_slots, void *stream);

/* 8-bit RGBA PNG of a host RGBA32F image as SavePngImage writes it
 * (src/FileIO.cpp:167-194): the extension becomes ".png", parent folders
 * are created, channels are clamp(v, 0, 1) * 255 + 0.5 truncated, row 0 of
 * the image is the bottom row of the file. */
int ymrt_save_png(const char *path, const float *rgba, uint32_t width, uint32_t height);

/* Running mean over frames (Accumulate_image_path_tracing). */
int ymrt_accumulate(const float *noise_rgba, float *accum_rgba, uint32_t width,
                    uint32_t height, int frame_index, int max_frames, void *stream);
/* Exposure, ACES fit and gamma in place (Gamma_correction_common). */
int ymrt_tonemap(float *rgba, uint32_t width, uint32_t height, float gamma,
                 float exposure, void *stream);

/* Halton sample values for testing the sampler bit for bit:
 * out[i*dims + d] = value of dimension (dim0 + d) of the sample with global
 * Halton index indices[i] (dims 0 and 1 are the unpermuted base-2/base-3
 * pixel-offset radical inverses).  Device pointers. */
int ymrt_halton_values(ymrt_scene *scene, const uint64_t *indices, uint32_t n,
                       uint32_t dim0, uint32_t dims, float *out, void *stream);
/* Global Halton index of (px, py, sample) for an image of width x height
 * (HaltonEnumerator::GetIndex), computed on the device. */
int ymrt_halton_enumerate(uint32_t width, uint32_t height, const uint32_t *px,
                          const uint32_t *py, const uint32_t *sample, uint32_t n,
                          uint64_t *out, void *stream);

/* ----------------------------------------------------------- .ymtmp files
 * Scene files in YMRT's text format, parsed as SceneParser::parsing_scene
 * (src/core/YMParser.cpp:2330-2433) with the block readers 971-2231 and the
 * SceneModule containers they fill.  The result owns a ymrt_scene_desc for
 * ymrt_scene_create (set camera.aspect_ratio and halton_seed first: the file
 * has neither).  A block that a reader rejects (missing name, undeclared or
 * duplicate name, wrong index counts, missing inputs) is skipped and listed
 * in the warnings, as the reference prints and skips it; a statement that is
 * not Begin(Type) / End(Type) / attribute=value, a nested or unclosed block
 * or an attribute that the block type does not know fails the whole load.
 * ImageTexture reads 8-bit PNG from <dir>/texture_<stem>_<ext>/.  A Volume's
 * density_texture_index becomes ymrt_volume_desc.density_texture. */
typedef struct ymrt_ymtmp ymrt_ymtmp;

#define YMRT_YMTMP_TRANSFORM 0u
#define YMRT_YMTMP_MESH 1u
#define YMRT_YMTMP_SPHERE 2u
#define YMRT_YMTMP_MATERIAL 3u
#define YMRT_YMTMP_TEXTURE 4u
#define YMRT_YMTMP_INSTANCE 5u
#define YMRT_YMTMP_DISTANT_LIGHT 6u
#define YMRT_YMTMP_VOLUME 7u

/* SceneParser::parsing_scene: path must end in ".ymtmp". */
int ymrt_ymtmp_load(const char *path, ymrt_ymtmp **out);
/* The same from text in memory; scene_path (may be NULL) locates textures. */
int ymrt_ymtmp_parse(const char *text, size_t length, const char *scene_path,
                     ymrt_ymtmp **out);
/* The loaded scene; valid until ymrt_ymtmp_free. */
const ymrt_scene_desc *ymrt_ymtmp_scene(const ymrt_ymtmp *file);
uint32_t ymrt_ymtmp_warning_count(const ymrt_ymtmp *file);
const char *ymrt_ymtmp_warning(const ymrt_ymtmp *file, uint32_t i);
/* The name -> (kind, index) map of load_scene_asset (YMParser.cpp:2284);
 * index is YMRT_INVALID_INDEX for an ImageTexture whose file failed to load. */
int ymrt_ymtmp_find(const ymrt_ymtmp *file, const char *name, uint32_t *kind,
                    uint32_t *index);
/* SceneModule::transforms / i_transforms[index] (InitTransformMatrix,
 * src/Transform.h:36-60), column-major; either output may be NULL. */
int ymrt_ymtmp_transform(const ymrt_ymtmp *file, uint32_t index,
                         float *object_to_world, float *world_to_object);
int ymrt_ymtmp_free(ymrt_ymtmp *file);

/* Texture graph rooted at `texture` evaluated at n points (TextureEval,
 * src/TextureEval.cuh:415-481, with the node functions 162-404):
 * coords[i*8 .. i*8+7] = (s, t, p_world xyz, p_object xyz), out[i*3 .. i*3+2]
 * = rgb.  Device pointers. */
int ymrt_texture_eval(ymrt_scene *scene, uint32_t texture, const float *coords,
                      uint32_t n, float *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* YMRT_H */

/*
 * oracle.h -- CPU oracle for the YMRT wavefront path tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference's
 * algorithms (file:line citations at each function), used by tests/ to check
 * the HIP path and by bench.py as the timed CPU baseline.  The product never
 * includes, links or calls anything here.
 *
 * Scene input is the same POD descriptor the product's C ABI takes
 * (include/ymrt.h), so both sides see identical float inputs.
 *
 * What pins each part (tests/test_oracle_*.py, tests/test_ymtmp.py) without
 * trusting the oracle:
 *   primes, prime sums, MT19937, permutation table, radical inverses (exact
 *   rationals), Halton enumerator (brute force), PCG4d (published algorithm);
 *   triangle test (float64), BVH closest/any hit (brute force over all
 *   triangles), watertightness, analytic sphere (closed-form roots); sRGB
 *   decode (IEC 61966-2-1), bilinear image lookup (ramps), Fresnel (textbook
 *   complex-index formulas), edge tint (reflectivity identity), GGX D
 *   (normalisation), every DefaultMtl lobe (sample/eval consistency, sampling
 *   histogram), Henyey-Greenstein (moments, sampled density), homogeneous
 *   free flight (closed form), grid free flight and transmittance (constant
 *   grid == homogeneous), surface NEE with MIS (analytic irradiance of a
 *   square emitter, a sphere emitter, a distant-light cone), accumulate and
 *   ACES tonemap; Perlin noise (zero on the lattice, lattice gradients of
 *   Perlin 2002 through the textbook hash, quintic fade along edges,
 *   continuity), fBm / turbulence / marble / wood / wave / checkerboard /
 *   polka-dot nodes (sums of the noise primitive, sign patterns), texture
 *   graph order (nested graphs, empty inputs), camera texture-coordinate
 *   differentials (pixel footprint), normal and bump maps (closed-form
 *   irradiance of a tilted Lambert surface; flat maps change nothing);
 *   .ymtmp transforms (glm from the reference tree, bit for bit, in
 *   oracle/_ref), the .ymtmp fixture (values read off the file).
 *
 * Parity unpinned (checked only against the HIP path, no independent truth):
 *   indirect_light_clamp, Russian roulette and the environment background
 *   (all off in the measured workloads); next-event estimation at volume
 *   vertices (EvalVolumeShapeLight, EvalVolumeDistantLight) and the mixture
 *   of overlapping media; the medium-record stack (RayTransfer priorities,
 *   BoundaryTransitionBunch, camera medium records, nested-dielectric
 *   pass-through); the polka-dot centre jitter (only cell membership is
 *   pinned); the .ymtmp statement rules beyond the cases tests/test_ymtmp.py
 *   lists.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>
#include "../include/ymrt.h"
#include "or_math.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PRIME_COUNT 1000
#define OR_MAX_BOUNDARY_RECORD 6
#define OR_MAX_NEARBY (OR_MAX_BOUNDARY_RECORD + 2)
#define OR_MAX_RAY_DEPTH 64
#define OR_MAX_FREE_PATH_LENGTH 2048

/* ------------------------------------------------------------------ sampler */
typedef struct { uint32_t s[624]; int i; } or_mt19937;
void or_mt_seed(or_mt19937 *mt, uint32_t seed);
uint32_t or_mt_next(or_mt19937 *mt);

uint32_t or_prime(uint32_t i);
uint32_t or_prime_sum(uint32_t i);
uint32_t or_halton_table_size(void);
void or_halton_table(uint32_t seed, uint32_t *table);

typedef struct {
    float x_scale, y_scale;
    int m_x, m_y, exp2, exp3, increment;
} or_enumerator;
void or_enum_init(or_enumerator *e, int width, int height);
uint64_t or_enum_index(const or_enumerator *e, int px, int py, int sample);

float or_radical_inverse(uint64_t index, uint32_t dim, const uint32_t *table);
float or_radical_inverse2(uint64_t index);
float or_radical_inverse3(uint64_t index);
float or_halton_value(uint64_t index, uint32_t dim, const uint32_t *table);
void or_pcg4d(uint32_t v[4]);

#define OR_SAMPLER_PCG 0
#define OR_SAMPLER_HALTON 1
typedef struct {
    int type;
    uint64_t index;      /* Halton global sample index            */
    uint32_t dim;
    const uint32_t *table;
    uint32_t pixel, sample, frame; /* PCG state                   */
} or_sampler;
float or_sampler_next(or_sampler *s);
extern float g_probe_u;

/* ----------------------------------------------------------------- geometry */
typedef struct { v3 pmax, pmin; } or_bbox; /* BBox3 stores p_max first */

typedef struct {
    or_bbox bbox;
    uint32_t a, b; /* internal: a = EMPTY, b = right child; leaf: a = offset, b = count */
} or_bnode;

typedef struct {
    uint32_t tri_count;
    const uint32_t *pidx, *nidx, *tidx;
    const float *pos, *nrm, *uv;
    uint32_t *order;        /* leaf order -> original triangle index */
    or_bnode *nodes;
    uint32_t node_count;
    or_bbox bounds;
} or_mesh;

typedef struct {
    float light_color[3]; float intensity;
} or_light_mtl;

typedef struct {
    uint32_t instance, triangle;
    float power, area;
} or_shape_light;

typedef struct {
    uint32_t width, height, channels;
    int wrap, color_space;
    float su, sv, ou, ov;
    float *texels;          /* width*height*channels decoded floats */
    const ymrt_texture_desc *desc;   /* type, constants, inputs, parameters */
} or_texture;

typedef struct {
    v2 st;
    v3 p_world, p_object;
} or_tex_coord;

typedef struct {
    uint32_t nx, ny, nz;
    const float *density;
    float max_density;
} or_grid;

typedef struct {
    uint32_t prim;
    float ior;
    uint32_t priority;
    int vol;
} or_record;

typedef struct {
    int max_idx;
    int count;
    or_record rec[OR_MAX_BOUNDARY_RECORD];
} or_transfer;

typedef struct {
    uint32_t mesh_count;
    or_mesh *meshes;
    uint32_t sphere_count;
    const ymrt_sphere_desc *spheres;
    uint32_t instance_count;
    const ymrt_instance_desc *instances; /* original order */
    uint32_t *inst_order;   /* top-leaf order -> original instance index */
    or_bnode *top;
    uint32_t top_count;
    uint32_t material_count;
    const ymrt_material_desc *materials;
    uint32_t texture_count;
    or_texture *textures;
    uint32_t grid_count;
    or_grid *grids;
    uint32_t volume_count;
    const ymrt_volume_desc *volumes;
    uint32_t distant_count;
    const ymrt_distant_light_desc *distant;
    float *distant_cos_max;   /* cos(radians(theta_max)) per distant light */
    uint32_t light_count;
    or_shape_light *lights;
    float *light_cdf;
    uint32_t *halton;
    ymrt_camera_desc camera;
    v3 cam_u, cam_v, cam_w;
    float tan_half_fov;
    or_transfer camera_transfer;
} or_scene;

typedef struct {
    v3 o, d;
    float t;
    uint32_t last_inst, last_tri;
} or_ray;

typedef struct {
    v3 bary;        /* (b0, b1, b2); for a volume hit the scatter position */
    float t;
    uint32_t inst;  /* original instance index or EMPTY */
    uint32_t tri;   /* original triangle index */
    int back;
    int volume;
} or_hit;

typedef struct {
    uint32_t prim;
    float t;
    int back;
} or_nearby;

or_scene *or_scene_create(const ymrt_scene_desc *desc);
void or_scene_destroy(or_scene *sc);

int or_trace_ray(const or_scene *sc, or_ray *ray, or_hit *hit, int *nearby_count, or_nearby *nearby);
int or_trace_shadow(const or_scene *sc, or_ray *ray);
int or_sphere_any(float radius, or_ray *object_ray);
int or_trace_volume(const or_scene *sc, or_ray *ray, or_hit *hit, int *nearby_count, or_nearby *nearby);
int or_intersect_triangle(const v3 p0, const v3 p1, const v3 p2, const or_ray *ray,
                          float *t, v3 *bary, int *back);
void or_fetch_geometry_normal(const or_scene *sc, const or_hit *h, v3 *p, v3 *perr, v3 *ng);
v3 or_offset_ray_origin(v3 p, v3 perr, v3 dir, v3 ng);

/* batch helpers over the C ABI ray layout */
void or_trace_closest_batch(const or_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits);
void or_trace_any_batch(const or_scene *sc, const ymrt_ray *rays, uint32_t n, uint32_t *occluded);
void or_brute_closest_batch(const or_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits);

/* ------------------------------------------------------------------ shading */
/* Renders rows [row0, row1) of one frame into out_rgba (width*height*4,
 * only those rows written): per pixel sum over samples of radiance/spp and
 * alpha = 1, i.e. the reference's noise image. */
void or_render_rows(const or_scene *sc, const ymrt_render_settings *rs, int frame,
                    uint32_t row0, uint32_t row1, float *out_rgba, uint64_t *ray_counts);
/* One path; returns its radiance (before the 1/spp scale). */
v3 or_trace_path(const or_scene *sc, const ymrt_render_settings *rs, int frame,
                 uint32_t px, uint32_t py, uint32_t sample, uint64_t *ray_counts);

/* BSDF probes for closed-form tests: fills a DefaultMtl BSDF for a hit whose
 * shading normal is +z, incoming direction wo (local, toward the viewer). */
int or_bsdf_eval(const ymrt_material_desc *m, v3 wo, v3 wi, float ray_ior, float ex_ior,
                 int hit_back, float *f_rgb, float *pdf);
int or_bsdf_sample(const ymrt_material_desc *m, v3 wo, float ray_ior, float ex_ior,
                   int hit_back, float u0, float u1, float u2, float *w_rgb, float *wi, float *pdf);
float or_hg_eval(float g, v3 wo, v3 wi, float *pdf);
int or_hg_sample(float g, float u0, float u1, v3 wo, v3 *wi, float *pdf);
v3 or_texture_eval(const or_scene *sc, uint32_t tex, float u, float v);
void or_first_hits(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t *ids);
int or_lambert_only(const ymrt_material_desc *m);

/* Parameters of the path-replay adjoint (or_render_adjoint); the gradient is
 * one float64 vector: albedo (3 per listed material), roughness (1 per listed
 * material), the texels of the target texture (width * height * channels of
 * decoded linear values), the voxels of the target volume's grid. */
typedef struct {
    int32_t *alb_slot;          /* per material: position in the albedo list or -1 */
    int32_t *rough_slot;        /* per material: position in the roughness list or -1 */
    uint32_t n_alb, n_rough;
    uint32_t tex;               /* target texture or 0xFFFFFFFF */
    uint32_t vol;               /* target volume or 0xFFFFFFFF */
    size_t off_rough, off_tex, off_vox, n_params;
} or_adj_targets;

/* Checks the targets against the scene (albedo: Lambert-only materials,
 * constant albedo; texture: an image used only as the diffuse albedo of
 * Lambert-only materials; roughness: metal-only materials, isotropic, no
 * textures; volume: a grid density with gray sigma_t) and lays out the
 * gradient vector.  0 on success, else the position of the first bad target
 * + 1 in the order albedo list, roughness list, texture, volume. */
int or_adjoint_prepare(const or_scene *sc, const uint32_t *alb, uint32_t n_alb, const uint32_t *rough,
                       uint32_t n_rough, uint32_t tex, uint32_t vol, or_adj_targets *tg);
void or_adjoint_free(or_adj_targets *tg);
void or_render_adjoint(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t row0, uint32_t row1,
                       const float *adj, const or_adj_targets *tg, double *grad, float *out);
void or_primary_rays(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t sample, ymrt_ray *out);
/* TextureEval over the texture graph rooted at tex */
v3 or_texture_graph(const or_scene *sc, uint32_t tex, or_tex_coord tc);
float or_perlin(v3 p);
float or_srgb_to_linear(float s);

void or_accumulate(const float *noise, float *accum, uint32_t w, uint32_t h, int frame);
void or_tonemap(float *rgba, uint32_t w, uint32_t h, float gamma, float exposure);

#ifdef __cplusplus
}
#endif
#endif

// dev_scene.hpp -- HBM layout of a scene as the gfx950 kernels read it.
//
// The reference keeps a binary BVH per mesh (32-B nodes, leaf size 1,
// src/BottomBVH.h:22-30) and fetches triangle corners through two levels of
// index indirection (src/ShadingUtilities.cuh:62-68).  Here:
//   * one array of 4-wide nodes (WNode, 128 B = one L2 line) holds every
//     bottom-level tree and the top-level tree;
//   * traversal triangles are pre-gathered in leaf order, 48 B each (three
//     float4: corner xyz + original triangle index in v0.w);
//   * shading triangles (positions, normals, uvs of the three corners) are
//     pre-gathered by original triangle index, 96 B each, read once per hit.
#pragma once

#include "dev_math.hpp"

namespace ym {

// Shading classes of a material (DevScene::mat_class): the lobes its
// DefaultMtl weighting leaves, known at compile time in a shading variant.
constexpr uint32_t kMatGeneral = 0, kMatDiffuse = 1, kMatMetal = 2;

// Child encoding in WNode::child and on the traversal stack (32 bits):
//   bit 31 clear        internal node index
//   bit 31 set          leaf: bits 30..3 = first primitive, bits 2..0 = count-1
//   0xFFFFFFFF          instance sentinel on the stack / empty child
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kSentinel = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeafPrims = 8;

YM_HD uint32_t leaf_code(uint32_t first, uint32_t count) { return kLeafBit | (first << 3) | (count - 1u); }
YM_HD uint32_t leaf_first(uint32_t code) { return (code & 0x7FFFFFFFu) >> 3; }
YM_HD uint32_t leaf_count(uint32_t code) { return (code & 7u) + 1u; }

struct alignas(128) WNode {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(WNode) == 128, "WNode must be one 128-B line");

struct alignas(16) DevInstance {  // 128 B
    Affine w2o;        // world -> object
    Affine o2w;        // object -> world
    uint32_t blas_root;
    uint32_t mesh;
    uint32_t material;
    int32_t inner_volume;
    uint32_t flags;    // bit 0: treat_as_boundary
    uint32_t shade_base;   // first shading triangle of the mesh
    uint32_t tri_count;
    float radius;          // sphere instances
};
static_assert(sizeof(DevInstance) == 128, "DevInstance size");

constexpr uint32_t kInstBoundary = 1u;
constexpr uint32_t kInstNoNormals = 2u;   // mesh without normals: shade with the face normal
constexpr uint32_t kInstSphere = 4u;      // analytic sphere of DevInstance::radius, no bottom-level tree

// Halton dimension d (base = the d-th prime, 0-based): its digit
// permutation at halton[perm_off] and 1 / base rounded to double.  One 16-B
// load gives the sampler all it needs before the digit loads.
struct alignas(16) HaltonDim {
    uint32_t base, perm_off;
    double inv_base;
};

struct alignas(16) ShadeTri {      // 96 B, by original triangle index
    float p[9];                    // p0, p1, p2
    float n[9];                    // n0, n1, n2 (face normal repeated when the mesh has none)
    float uv[6];                   // uv0, uv1, uv2 (reference defaults when absent)
};
static_assert(sizeof(ShadeTri) == 96, "ShadeTri size");

struct DevMaterial {               // ymrt_material_desc, 136 B
    uint32_t type;
    float diffuse_albedo[3]; uint32_t diffuse_albedo_tex;
    float specular_albedo[3]; uint32_t specular_albedo_tex;
    float roughness_x; uint32_t roughness_x_tex;
    float roughness_y; uint32_t roughness_y_tex;
    float specular_weight; uint32_t specular_weight_tex;
    float metalness; uint32_t metalness_tex;
    float transmission_weight; uint32_t transmission_weight_tex;
    float ior_n; uint32_t ior_priority;
    float coat_albedo[3]; uint32_t coat_albedo_tex;
    float coat_weight; uint32_t coat_weight_tex;
    float coat_thickness; uint32_t coat_thickness_tex;
    float coat_roughness_x; uint32_t coat_roughness_x_tex;
    float coat_roughness_y; uint32_t coat_roughness_y_tex;
    float coat_ior;
    uint32_t normal_mapping_tex;
    uint32_t bump_mapping_tex;
    float light_color[3];
    float intensity;
};

struct DevTexture {
    uint32_t width, height, channels;
    int32_t wrap, color_space;
    float su, sv, ou, ov;
    uint64_t offset;               // first texel (uint32 RGBA8 words) in the texel pool
    uint32_t type;                 // YMRT_TEXTURE_*
    float value[3];
    float frequency, lacunarity, gain, amplitude, offset_f, variation, radius;
    int32_t layers, normalized, axes;
    float frequency1, lacunarity1, gain1;
    int32_t layers1;
    uint32_t prog_start, prog_len; // post-order program of the graph rooted here
    uint32_t f32;                  // IMAGE: texels are linear floats (offset into texels_f32)
};

// Texture program codes: a texture index, or kTexBlack for an unconnected input.
constexpr uint32_t kTexBlack = 0xFFFFFFFEu;
constexpr int kTexStack = 8;       // TextureEval's texture_stack_size

struct DevGrid {
    uint32_t nx, ny, nz;
    float max_density;
    uint64_t offset;               // first voxel in the density pool
};

struct DevVolume {
    float sigma_t[3];
    float albedo[3];
    float g;
    int32_t grid;                  // voxel-grid density, or -1
    Affine w2o;
    int32_t tex;                   // texture-graph density (x channel), or -1
    float tex_bound;               // its majorant, Texture::GetUpperBound
};

// A shape light as the scene build finds it (LoadShapeLight): power for the
// selection CDF, and the sampling area.
struct DevLight {
    uint32_t instance, triangle;
    float power, area;
};

// A shape light with everything next-event estimation reads from it: the
// emitting triangle in object space with its vertex normals (or the sphere's
// radius), the instance transforms and the emitted radiance
// light_color * intensity.  Sampling a light is then one dependent load
// after the pick, where the scene's own records take three (light ->
// instance -> triangle and material).
struct alignas(16) DevLightRec {
    Affine w2o, o2w;
    float p[9], n[9];
    float le[3];
    float area;         // DevLight::area
    float radius;       // sphere lights
    uint32_t triangle;  // kEmpty: a sphere
};
static_assert(sizeof(DevLightRec) == 192, "DevLightRec size");

// Light-selection CDFs of up to this many lights are also kept in the scene
// record, so the pick reads them with scalar loads.
constexpr uint32_t kLightCdfInline = 8;

struct DevDistant {
    f3 color;                      // light_color
    float intensity;
    float cos_max;                 // cos(radians(theta_max))
    Affine o2w, w2o;               // the light's frame
};

struct CameraRecord {
    uint32_t prim;
    float ior;
    uint32_t priority;
    int32_t vol;
};

struct DevCamera {
    f3 position, u, v, w;
    float tan_half_fov, aspect;
    int32_t record_count, max_idx;
    CameraRecord rec[kMaxBoundaryRecord];
};

// Pointers and counts handed to every kernel by value.
struct DevScene {
    const WNode *nodes;
    const float4 *tris;            // 3 float4 per traversal triangle
    const DevInstance *instances;
    const ShadeTri *shade;
    const DevMaterial *materials;
    const DevTexture *textures;
    const uint32_t *tex_prog;      // texture programs (codes)
    const uint32_t *texels;
    const float *texels_f32;       // float image textures, channels per texel
    const float *tex_lut;          // 2 x 256 decode tables (sRGB, linear)
    const DevGrid *grids;
    const float *voxels;
    const DevVolume *volumes;
    const DevLightRec *light_recs;  // per shape light
    const float *light_cdf;
    float light_cdf_inline[kLightCdfInline];  // light_cdf when light_count <= kLightCdfInline
    const DevDistant *distant;
    const uint32_t *halton;        // permutation table
    const HaltonDim *halton_dims;  // per dimension: prime, permutation offset, 1 / prime
    uint32_t tlas_root;            // all instances: node index, leaf code, or kSentinel
    uint32_t tlas_root_solid;      // instances that are not medium boundaries (shadow rays)
    uint32_t tlas_root_boundary;   // medium boundaries only (transmittance walk)
    uint32_t instance_count, material_count, texture_count, grid_count, volume_count, light_count;
    uint32_t distant_count;
    uint32_t has_boundaries;       // any instance with treat_as_boundary
    uint32_t has_spheres;          // any sphere instance
    uint32_t has_graph;            // procedural texture nodes or normal / bump maps
    uint32_t has_tex_media;        // any volume whose density is a texture graph
    const uint32_t *mat_class;     // per material: kMatDiffuse, kMatMetal or 0 (shading variant)
    DevCamera camera;
};

}  // namespace ym

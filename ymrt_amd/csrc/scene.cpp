// scene.cpp -- host side of ymrt_scene_create: builds the wide BVHs, the
// shading/light/texture tables and the sampler table, computes the camera's
// starting media, and uploads everything to HBM.
//
// Replaces the host steps of GuiModule::UpdateScene (src/core/GuiModule.cpp:116-271):
// TopBVHBuilder::BuildSceneBVH, SceneModule::LoadShapeLight
// (src/core/SceneModule.cpp:1203-1304), InitHaltonPermuteTable (1412-1432),
// Camera::InitCameraRayTransfer (src/Camera.cpp:15-78), the editor->render
// camera copy with the image aspect ratio, and the UPLOAD_TO_GPU calls.
#include <chrono>
#include <functional>
#include <cmath>
#include <cstring>
#include <random>

#include "bvh.hpp"
#include "dev_traverse.hpp"
#include "internal.hpp"

namespace ym {

namespace {

template <typename T>
int upload(ymrt_scene *sc, const std::vector<T> &v, const T **out)
{
    size_t bytes = sizeof(T) * std::max<size_t>(v.size(), 1);
    void *p = nullptr;
    YM_HIP(hipMalloc(&p, bytes));
    sc->allocations.push_back(p);
    if (!v.empty()) YM_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    sc->info.device_bytes += bytes;
    *out = static_cast<const T *>(p);
    return YMRT_OK;
}

Affine affine_from_colmajor(const float *m)
{
    Affine a;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) a.r[r * 4 + c] = m[c * 4 + r];
    return a;
}

f3 ld3(const float *a, uint32_t i) { return f3{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

// RayTransfer operations for the camera records (src/Ray.h:131-219)
struct Transfer {
    int max_idx = -1, count = 0;
    CameraRecord rec[kMaxBoundaryRecord];
    bool push(uint32_t prim, float ior, uint32_t prio, int vol)
    {
        if (count == kMaxBoundaryRecord) return false;
        rec[count] = CameraRecord{prim, ior, prio, vol};
        ++count;
        if (count == 1) {
            max_idx = 0;
            return true;
        }
        if (prio >= rec[max_idx].priority) max_idx = count - 1;
        return true;
    }
};

// Closest-hit traversals keep the nearby-hit list when medium boundaries,
// dielectrics or IOR priorities can change a path's medium records.
bool needs_nearby(bool has_boundaries, const std::vector<DevMaterial> &materials)
{
    bool r = has_boundaries;
    for (const DevMaterial &m : materials)
        if (m.type == YMRT_MATERIAL_DEFAULT &&
            (m.transmission_weight > 0.0f || m.transmission_weight_tex != kEmpty || m.ior_priority != 0))
            r = true;
    return r;
}

// A DefaultMtl whose coat, metal, specular and transmission weights are 0
// (untextured), with no normal or bump map and a coat that cannot overflow
// (albedo in [0, 1], finite thickness): InitBSDFSettings then leaves the
// diffuse lobe alone, with the albedo itself as its colour, bit for bit, so
// a shading variant with only that lobe gives the same floats.
uint32_t diffuse_only(const DevMaterial &m)
{
    bool ok = m.type == YMRT_MATERIAL_DEFAULT && m.coat_weight == 0.0f && m.coat_weight_tex == kEmpty &&
              m.metalness == 0.0f && m.metalness_tex == kEmpty && m.specular_weight == 0.0f &&
              m.specular_weight_tex == kEmpty && m.transmission_weight == 0.0f &&
              m.transmission_weight_tex == kEmpty && m.coat_albedo_tex == kEmpty &&
              m.coat_thickness_tex == kEmpty && m.normal_mapping_tex == kEmpty && m.bump_mapping_tex == kEmpty &&
              std::isfinite(m.coat_thickness) && m.coat_thickness >= 0.0f && std::isfinite(m.ior_n);
    for (int c = 0; c < 3; ++c) ok = ok && m.coat_albedo[c] >= 0.0f && m.coat_albedo[c] <= 1.0f;
    return ok ? 1u : 0u;
}

// A DefaultMtl with metalness exactly 1 (untextured), zero coat weight, a
// coat whose transmittance stays finite (albedo in [0, 1], thickness in
// [0, 1e6]: -log(1e-8) * 1e6 / cos >= -1e37), finite untextured specular and transmission
// weights and no normal or bump map: InitBSDFSettings then leaves the metal
// lobe with colour (1, 1, 1) and a diffuse lobe of weight 0 (the
// (1 - metalness) factor), bit for bit.
uint32_t metal_only(const DevMaterial &m)
{
    bool ok = m.type == YMRT_MATERIAL_DEFAULT && m.metalness == 1.0f && m.metalness_tex == kEmpty &&
              m.coat_weight == 0.0f && m.coat_weight_tex == kEmpty && m.coat_albedo_tex == kEmpty &&
              m.coat_thickness_tex == kEmpty && m.specular_weight_tex == kEmpty &&
              m.transmission_weight_tex == kEmpty && m.normal_mapping_tex == kEmpty && m.bump_mapping_tex == kEmpty &&
              m.coat_thickness >= 0.0f && m.coat_thickness <= 1e6f && std::isfinite(m.ior_n) &&
              std::isfinite(m.specular_weight) && std::isfinite(m.transmission_weight);
    for (int c = 0; c < 3; ++c) ok = ok && m.coat_albedo[c] >= 0.0f && m.coat_albedo[c] <= 1.0f;
    return ok ? 1u : 0u;
}

uint32_t shading_class(const DevMaterial &m)
{
    return diffuse_only(m) ? kMatDiffuse : metal_only(m) ? kMatMetal : kMatGeneral;
}

// Texture slots of a material, all of which must name a texture or be empty.
bool material_textures_valid(const DevMaterial &m, uint32_t texture_count)
{
    const uint32_t ids[] = {m.diffuse_albedo_tex, m.specular_albedo_tex, m.roughness_x_tex, m.roughness_y_tex,
                            m.specular_weight_tex, m.metalness_tex, m.transmission_weight_tex,
                            m.coat_albedo_tex, m.coat_weight_tex, m.coat_thickness_tex,
                            m.coat_roughness_x_tex, m.coat_roughness_y_tex, m.normal_mapping_tex,
                            m.bump_mapping_tex};
    for (uint32_t t : ids)
        if (t != kEmpty && t >= texture_count) return false;
    return true;
}

bool uses_texture_graph(bool procedural_textures, const std::vector<DevMaterial> &materials)
{
    bool r = procedural_textures;
    for (const DevMaterial &m : materials)
        if (m.normal_mapping_tex != kEmpty || m.bump_mapping_tex != kEmpty) r = true;
    return r;
}

float fetch_ior(const DevMaterial &m, uint32_t *prio)
{
    if (m.type == YMRT_MATERIAL_DEFAULT) {
        *prio = m.ior_priority;
        return m.ior_n;
    }
    *prio = 0;
    return 1.0f;
}

}  // namespace

// Host copies kept while the scene is built (the camera walk traverses them).
struct HostScene {
    std::vector<WNode> nodes;
    std::vector<float4> tris;
    std::vector<DevInstance> instances;
    std::vector<ShadeTri> shade;
    std::vector<DevMaterial> materials;
};

// Texture::GetUpperBound (src/Texture.h:1181-1251): the majorant of a
// density texture, per node type of the graph's root (images 0, so a medium
// whose only density is an image never scatters; fBm and turbulence the sum
// of amplitude + offset over their layers, in float as written there).
static float texture_upper_bound(const ymrt_texture_desc &t)
{
    switch (t.type) {
    case YMRT_TEXTURE_IMAGE:
        return 0.0f;
    case YMRT_TEXTURE_CONSTANT_FLOAT:
        return t.value[0];
    case YMRT_TEXTURE_CONSTANT_RGB:
        return std::max(t.value[0], std::max(t.value[1], t.value[2]));
    case YMRT_TEXTURE_FBM:
    case YMRT_TEXTURE_TURBULENCE: {
        float r = 0.0f, amp = t.amplitude, off = t.offset;
        for (int i = 0; i < t.layer_count; ++i) {
            r += amp + off;
            amp *= t.gain;
            off *= t.gain;
        }
        return r;
    }
    case YMRT_TEXTURE_WAVE: {
        // SafeRcp (src/MathCommon.h:103-106)
        auto safe_rcp = [](float f) { return std::fabs(f) < 1E-30f ? (f > 0.0f ? 1E30f : -1E30f) : 1.0f / f; };
        return safe_rcp(1.0f - t.gain) * safe_rcp(1.0f - t.gain1);
    }
    default:  // checkerboard, noise, marble, wood, polka dot
        return 1.0f;
    }
}

static void camera_setup(const ymrt_scene_desc *d, DevCamera &cam)
{
    // Camera::SetDir / MakeCameraBasis (src/Camera.h:28-32, 311-322)
    f3 w{d->camera.w[0], d->camera.w[1], d->camera.w[2]};
    w = w * safe_rcp(length(w));
    const f3 up{0.0f, 1.0f, 0.0f};
    double mu[3], mv[3];
    mu[0] = (double)up.y * (double)w.z - (double)up.z * (double)w.y;
    mu[1] = -((double)up.x * (double)w.z - (double)up.z * (double)w.x);
    mu[2] = (double)up.x * (double)w.y - (double)up.y * (double)w.x;
    f3 u = f3{(float)mu[0], (float)mu[1], (float)mu[2]} *
           safe_rcp((float)std::sqrt(mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2]));
    mv[0] = (double)w.y * (double)u.z - (double)w.z * (double)u.y;
    mv[1] = -((double)w.x * (double)u.z - (double)w.z * (double)u.x);
    mv[2] = (double)w.x * (double)u.y - (double)w.y * (double)u.x;
    f3 v = f3{(float)mv[0], (float)mv[1], (float)mv[2]} *
           safe_rcp((float)std::sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]));
    cam.position = f3{d->camera.position[0], d->camera.position[1], d->camera.position[2]};
    cam.u = u;
    cam.v = v;
    cam.w = w;
    cam.tan_half_fov = tanf(0.5f * d->camera.fov);
    cam.aspect = d->camera.aspect_ratio;
}

// Camera::GenerateRay at the image centre (src/Camera.h:164-170)
static Query center_ray(const DevCamera &cam)
{
    float height = 2.0f * cam.tan_half_fov;
    float width = height * cam.aspect;
    f3 dir = -cam.w + ((0.5f - 0.5f) * width) * cam.u + ((0.5f - 0.5f) * height) * cam.v;
    Query q;
    q.o = cam.position;
    q.d = normalize(dir);
    q.tmin = kTMin;
    q.tmax = kTMax;
    q.last_inst = kEmpty;
    q.last_tri = kEmpty;
    return q;
}

// Camera::InitCameraRayTransfer (src/Camera.cpp:15-78).  Instances are visited
// in the order the centre ray first meets them (the reference iterates an
// unordered_map).
static void camera_media(const DevScene &hs, const HostScene &h, const ymrt_scene_desc *d, DevCamera &cam,
                         uint32_t stack_slots)
{
    Transfer t;
    std::vector<uint32_t> seen;
    Query q = center_ray(cam);
    std::vector<uint32_t> stack(1024);
    (void)stack_slots;
    uint32_t nbp[kNearbyCap];
    float nbt[kNearbyCap];
    int nbb[kNearbyCap];
    LaneScratch<1> ls{stack.data(), nbp, nbt, nbb, 0, nullptr, 0, 0};
    for (int depth = 0; depth < 48; ++depth) {
        Hit hit{};
        int nbc = 0;
        TraceCounters cnt{0, 0};
        int r = traverse<kModeAll, false, false, false, true, 1024>(hs, q, hit, nbc, ls, cnt);
        if (r != 1) break;
        f3 p, perr, ng;
        hit_geometry(hs, hit.inst, hit.tri, hit.b0, hit.b1, hit.b2, p, perr, ng);
        if (std::find(seen.begin(), seen.end(), hit.inst) == seen.end()) seen.push_back(hit.inst);
        q.o = offset_ray_origin(p, perr, q.d, ng);
        q.last_inst = hit.inst;
        q.last_tri = hit.tri;
    }
    for (uint32_t ii : seen) {
        const DevInstance &in = h.instances[ii];
        Query c = center_ray(cam);
        f3 po = xf_point(in.w2o, c.o), pd = xf_dir(in.w2o, c.d);
        RayPre pos = make_pre(po, pd), neg = make_pre(po, -pd);
        bool hp = false, hn = false;
        if (in.flags & kInstSphere) {
            SphereTest st;
            hp = sphere_test(po, pd, in.radius, kTMin, kTMax, st);
            hn = sphere_test(po, -pd, in.radius, kTMin, kTMax, st);
        }
        for (uint32_t k = 0; k < in.tri_count && !(hp && hn); ++k) {
            const ShadeTri &st = h.shade[in.shade_base + k];
            f3 p0{st.p[0], st.p[1], st.p[2]}, p1{st.p[3], st.p[4], st.p[5]}, p2{st.p[6], st.p[7], st.p[8]};
            TriHit th;
            if (!hp && tri_test(pos, p0, p1, p2, th) && th.t >= kTMin && th.t <= kTMax) hp = true;
            if (!hn && tri_test(neg, p0, p1, p2, th) && th.t >= kTMin && th.t <= kTMax) hn = true;
        }
        if (hp && hn) {
            uint32_t prio;
            float ior = fetch_ior(h.materials[in.material], &prio);
            if (!t.push(ii, ior, prio, in.inner_volume)) break;
        }
    }
    cam.record_count = t.count;
    cam.max_idx = t.max_idx;
    for (int i = 0; i < kMaxBoundaryRecord; ++i) cam.rec[i] = t.rec[i];
    (void)d;
}

// light_color * intensity, as the shading code forms a light's radiance
static void light_emission_of(const DevMaterial &m, float le[3])
{
    for (int c = 0; c < 3; ++c) le[c] = m.light_color[c] * m.intensity;
}

int build_scene(ymrt_scene *sc, const ymrt_scene_desc *d)
{
    auto t0 = std::chrono::high_resolution_clock::now();
    YM_REQUIRE(d != nullptr, "scene desc is NULL");
    YM_REQUIRE(d->mesh_count == 0 || d->meshes != nullptr, "meshes is NULL");
    YM_REQUIRE(d->instance_count == 0 || d->instances != nullptr, "instances is NULL");
    YM_REQUIRE(d->material_count == 0 || d->materials != nullptr, "materials is NULL");
    static_assert(sizeof(DevMaterial) == sizeof(ymrt_material_desc), "material layout");

    HostScene h;
    // ---------------------------------------------------------------- meshes
    std::vector<uint32_t> blas_root(d->mesh_count), shade_base(d->mesh_count), mesh_flags(d->mesh_count, 0);
    uint32_t max_blas_depth = 0;
    uint64_t total_tris = 0;
    for (uint32_t mi = 0; mi < d->mesh_count; ++mi) {
        const ymrt_mesh_desc &m = d->meshes[mi];
        YM_REQUIRE(m.triangle_count == 0 || (m.position_indices && m.positions), "mesh without positions");
        for (uint64_t i = 0; i < 3ull * m.triangle_count; ++i)
            YM_REQUIRE(m.position_indices[i] < m.position_count, "position index out of range");
        const bool has_n = m.normal_indices && m.normals;
        const bool has_uv = m.texcoord_indices && m.texcoords;
        if (has_n)
            for (uint64_t i = 0; i < 3ull * m.triangle_count; ++i)
                YM_REQUIRE(m.normal_indices[i] < m.normal_count, "normal index out of range");
        if (has_uv)
            for (uint64_t i = 0; i < 3ull * m.triangle_count; ++i)
                YM_REQUIRE(m.texcoord_indices[i] < m.texcoord_count, "texcoord index out of range");
        if (!has_n) mesh_flags[mi] |= kInstNoNormals;
        shade_base[mi] = (uint32_t)h.shade.size();
        std::vector<BuildPrim> prims(m.triangle_count);
        for (uint32_t t = 0; t < m.triangle_count; ++t) {
            f3 p[3];
            ShadeTri st;
            for (int k = 0; k < 3; ++k) {
                p[k] = ld3(m.positions, m.position_indices[3 * t + k]);
                st.p[3 * k] = p[k].x;
                st.p[3 * k + 1] = p[k].y;
                st.p[3 * k + 2] = p[k].z;
            }
            f3 g = cross(p[1] - p[0], p[2] - p[0]);
            for (int k = 0; k < 3; ++k) {
                f3 n = has_n ? ld3(m.normals, m.normal_indices[3 * t + k]) : g;
                st.n[3 * k] = n.x;
                st.n[3 * k + 1] = n.y;
                st.n[3 * k + 2] = n.z;
            }
            static const float kDefUv[6] = {0.0f, 0.0f, 1.0f, 0.0f, 1.0f, 1.0f};
            for (int k = 0; k < 3; ++k) {
                if (has_uv) {
                    uint32_t ti = m.texcoord_indices[3 * t + k];
                    st.uv[2 * k] = m.texcoords[2 * ti];
                    st.uv[2 * k + 1] = m.texcoords[2 * ti + 1];
                } else {
                    st.uv[2 * k] = kDefUv[2 * k];
                    st.uv[2 * k + 1] = kDefUv[2 * k + 1];
                }
            }
            h.shade.push_back(st);
            BuildPrim &bp = prims[t];
            for (int k = 0; k < 3; ++k) {
                const float pk[3] = {p[k].x, p[k].y, p[k].z};
                bp.box.grow(pk);
            }
            bp.box.fix();
            for (int a = 0; a < 3; ++a) bp.c[a] = 0.5f * (bp.box.lo[a] + bp.box.hi[a]);
            bp.id = t;
        }
        uint32_t tri_base = (uint32_t)(h.tris.size() / 3);
        YM_REQUIRE((uint64_t)tri_base + m.triangle_count < (1ull << 28), "too many triangles for leaf encoding");
        BinaryBuilder bb(prims, 4);
        std::vector<BinNode> bin = bb.build();
        uint32_t depth = 0;
        blas_root[mi] = collapse_wide(bin, h.nodes, tri_base, &depth);
        max_blas_depth = std::max(max_blas_depth, depth);
        for (const BuildPrim &bp : prims) {
            const ShadeTri &st = h.shade[shade_base[mi] + bp.id];
            h.tris.push_back(make_float4(st.p[0], st.p[1], st.p[2], u2f(bp.id)));
            h.tris.push_back(make_float4(st.p[3], st.p[4], st.p[5], 0.0f));
            h.tris.push_back(make_float4(st.p[6], st.p[7], st.p[8], 0.0f));
        }
    }
    // ------------------------------------------------------------- materials
    for (uint32_t i = 0; i < d->material_count; ++i) {
        DevMaterial m;
        std::memcpy(&m, &d->materials[i], sizeof(m));
        h.materials.push_back(m);
    }
    // ------------------------------------------------------------- instances
    std::vector<BuildPrim> iprims;
    uint32_t has_boundaries = 0, has_spheres = 0;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const ymrt_instance_desc &id = d->instances[i];
        YM_REQUIRE(id.geometry_type == YMRT_GEOMETRY_MESH || id.geometry_type == YMRT_GEOMETRY_SPHERE,
                   "instance geometry_type must be MESH or SPHERE");
        const bool sphere = id.geometry_type == YMRT_GEOMETRY_SPHERE;
        YM_REQUIRE(sphere ? id.geometry < d->sphere_count : id.geometry < d->mesh_count,
                   "instance geometry index out of range");
        YM_REQUIRE(id.material < d->material_count, "instance material index out of range");
        YM_REQUIRE(id.inner_volume < 0 || (uint32_t)id.inner_volume < d->volume_count,
                   "instance inner_volume out of range");
        DevInstance in{};
        in.w2o = affine_from_colmajor(id.world_to_object);
        in.o2w = affine_from_colmajor(id.object_to_world);
        in.blas_root = sphere ? kSentinel : blas_root[id.geometry];
        in.mesh = id.geometry;
        in.material = id.material;
        in.inner_volume = id.inner_volume;
        in.flags = (id.treat_as_boundary ? kInstBoundary : 0u) | (sphere ? kInstSphere : mesh_flags[id.geometry]);
        in.shade_base = sphere ? 0u : shade_base[id.geometry];
        in.tri_count = sphere ? 0u : d->meshes[id.geometry].triangle_count;
        in.radius = sphere ? d->spheres[id.geometry].radius : 0.0f;
        has_boundaries |= id.treat_as_boundary ? 1u : 0u;
        has_spheres |= sphere ? 1u : 0u;
        total_tris += in.tri_count;
        h.instances.push_back(in);
        if (!sphere && in.tri_count == 0) continue;
        // world bounds: of the transformed vertices for a mesh (a rotated
        // mesh's transformed object box can be far larger, and every ray
        // through the excess enters the instance for nothing), of the
        // transformed object box for a sphere
        Aabb wb;
        if (sphere) {
            const float r = std::fabs(in.radius);
            for (int c = 0; c < 8; ++c) {
                f3 corner{(c & 1) ? r : -r, (c & 2) ? r : -r, (c & 4) ? r : -r};
                f3 w = xf_point(in.o2w, corner);
                const float wk[3] = {w.x, w.y, w.z};
                wb.grow(wk);
            }
        } else {
            const ymrt_mesh_desc &m = d->meshes[id.geometry];
            for (uint32_t k = 0; k < m.position_count; ++k) {
                f3 w = xf_point(in.o2w, f3{m.positions[3 * k], m.positions[3 * k + 1], m.positions[3 * k + 2]});
                const float wk[3] = {w.x, w.y, w.z};
                wb.grow(wk);
            }
        }
        // pad by a relative margin so rounding of the transform never culls
        for (int a = 0; a < 3; ++a) {
            float pad = 1e-5f * std::max({1.0f, std::fabs(wb.lo[a]), std::fabs(wb.hi[a])});
            wb.lo[a] -= pad;
            wb.hi[a] += pad;
        }
        wb.fix();
        BuildPrim bp;
        bp.box = wb;
        for (int a = 0; a < 3; ++a) bp.c[a] = 0.5f * (wb.lo[a] + wb.hi[a]);
        bp.id = i;
        iprims.push_back(bp);
    }
    // Three top-level trees: all instances (closest hit), non-boundary
    // instances (shadow any-hit), boundary instances (transmittance walk).
    uint32_t tlas_depth = 0;
    auto build_tlas = [&](int which) -> uint32_t {
        std::vector<BuildPrim> sel;
        for (const BuildPrim &bp : iprims) {
            const bool boundary = d->instances[bp.id].treat_as_boundary != 0;
            if (which == 0 || (which == 1 && !boundary) || (which == 2 && boundary)) sel.push_back(bp);
        }
        if (sel.empty()) return kSentinel;
        BinaryBuilder tb(sel, 1);
        std::vector<BinNode> bin = tb.build();
        // leaves of size 1 referencing the original instance index
        for (BinNode &n : bin)
            if (n.left < 0) n.first = sel[n.first].id;
        uint32_t depth = 0;
        uint32_t root = collapse_wide(bin, h.nodes, 0, &depth);
        tlas_depth = std::max(tlas_depth, depth);
        return root;
    };
    const uint32_t tlas_root = build_tlas(0);
    const uint32_t tlas_root_solid = build_tlas(1);
    const uint32_t tlas_root_boundary = build_tlas(2);
    sc->max_stack = 3 * max_blas_depth + 3 * tlas_depth + 2;
    sc->has_boundaries = has_boundaries != 0;
    sc->needs_nearby = needs_nearby(sc->has_boundaries, h.materials);
    sc->info.triangle_count = total_tris;
    sc->info.node_count = h.nodes.size();

    DevScene &ds = sc->dev;
    ds = DevScene{};
    ds.tlas_root = tlas_root;
    ds.tlas_root_solid = tlas_root_solid;
    ds.tlas_root_boundary = tlas_root_boundary;
    ds.instance_count = (uint32_t)iprims.size() ? d->instance_count : 0;
    ds.material_count = d->material_count;
    ds.has_boundaries = has_boundaries;
    ds.has_spheres = has_spheres;
    // ---------------------------------------------------------- textures / grids
    std::vector<DevTexture> texs;
    std::vector<uint32_t> texels;
    std::vector<float> texels_f32;
    YM_REQUIRE(d->texture_count == 0 || d->textures != nullptr, "textures is NULL");
    for (uint32_t i = 0; i < d->texture_count; ++i) {
        const ymrt_texture_desc &t = d->textures[i];
        YM_REQUIRE(t.type <= YMRT_TEXTURE_WAVE, "unknown texture type");
        DevTexture dt{};
        dt.type = t.type;
        for (int k = 0; k < 3; ++k) dt.value[k] = t.value[k];
        dt.frequency = t.frequency;
        dt.lacunarity = t.lacunarity;
        dt.gain = t.gain;
        dt.amplitude = t.amplitude;
        dt.offset_f = t.offset;
        dt.variation = t.variation;
        dt.radius = t.radius;
        dt.layers = t.layer_count;
        dt.normalized = t.normalized;
        dt.axes = t.turbulence_axes;
        dt.frequency1 = t.frequency1;
        dt.lacunarity1 = t.lacunarity1;
        dt.gain1 = t.gain1;
        dt.layers1 = t.layer_count1;
        if (t.type != YMRT_TEXTURE_IMAGE) {
            dt.offset = texels.size();
            texs.push_back(dt);
            continue;
        }
        YM_REQUIRE(t.width > 0 && t.height > 0 && (t.texels || t.texels_f32), "empty texture");
        YM_REQUIRE(t.channels >= 1 && t.channels <= 4, "texture channels must be 1..4");
        dt.width = t.width;
        dt.height = t.height;
        dt.channels = t.channels;
        dt.wrap = t.wrap_mode;
        dt.color_space = t.color_space;
        dt.su = t.scale_u;
        dt.sv = t.scale_v;
        dt.ou = t.offset_u;
        dt.ov = t.offset_v;
        const size_t n = (size_t)t.width * t.height;
        if (t.texels_f32) {
            dt.f32 = 1;
            dt.offset = texels_f32.size();
            texels_f32.insert(texels_f32.end(), t.texels_f32, t.texels_f32 + n * t.channels);
            texs.push_back(dt);
            continue;
        }
        dt.offset = texels.size();
        for (size_t k = 0; k < n; ++k) {
            uint32_t w = 0;
            for (uint32_t c = 0; c < t.channels; ++c) w |= (uint32_t)t.texels[k * t.channels + c] << (8 * c);
            texels.push_back(w);
        }
        texs.push_back(dt);
    }
    // Texture programs: each graph flattened in post order (inputs black then
    // white before their node) for a value stack of at most kTexStack entries,
    // the bound of TextureEval's explicit stack (src/TextureEval.cuh:415-481).
    std::vector<uint32_t> tex_prog;
    for (uint32_t i = 0; i < d->texture_count; ++i) {
        const uint32_t start = (uint32_t)tex_prog.size();
        int depth = 0, max_depth = 0;
        bool ok = true;
        std::function<void(uint32_t, int)> emit = [&](uint32_t ti, int level) {
            if (!ok) return;
            if (level > kTexStack) {
                ok = false;
                return;
            }
            if (ti == kEmpty) {
                tex_prog.push_back(kTexBlack);
                max_depth = std::max(max_depth, ++depth);
                return;
            }
            const ymrt_texture_desc &t = d->textures[ti];
            if (t.type >= YMRT_TEXTURE_CHECKERBOARD) {
                for (int c = 0; c < 2; ++c) {
                    if (t.child[c] != kEmpty && t.child[c] >= d->texture_count) {
                        ok = false;
                        return;
                    }
                    emit(t.child[c], level + 1);
                }
                depth -= 2;
            }
            tex_prog.push_back(ti);
            max_depth = std::max(max_depth, ++depth);
        };
        emit(i, 1);
        YM_REQUIRE(ok, "texture graph has an input out of range, a cycle or more than 8 levels");
        YM_REQUIRE(max_depth <= kTexStack, "texture graph needs more than 8 stack entries");
        texs[i].prog_start = start;
        texs[i].prog_len = (uint32_t)tex_prog.size() - start;
    }
    for (uint32_t i = 0; i < d->material_count; ++i)
        YM_REQUIRE(material_textures_valid(h.materials[i], d->texture_count), "material texture index out of range");
    sc->procedural_textures = false;
    for (const DevTexture &t : texs)
        if (t.type >= YMRT_TEXTURE_CHECKERBOARD) sc->procedural_textures = true;
    ds.has_graph = uses_texture_graph(sc->procedural_textures, h.materials) ? 1 : 0;
    std::vector<float> lut(512);
    for (int i = 0; i < 256; ++i) {
        float v = (float)i * (1.0f / 255.0f);
        // SRGBToLinear (src/MathCommon.h:136-139): double pow with the float 2.4f
        lut[i] = v > 0.04045 ? (float)std::pow(((double)v + 0.055) / 1.055, (double)2.4f) : (v / 12.92f);
        lut[256 + i] = v;
    }
    std::vector<DevGrid> grids;
    std::vector<float> voxels;
    for (uint32_t i = 0; i < d->grid_count; ++i) {
        const ymrt_grid_desc &g = d->grids[i];
        YM_REQUIRE(g.nx && g.ny && g.nz && g.density, "empty density grid");
        DevGrid dg;
        dg.nx = g.nx;
        dg.ny = g.ny;
        dg.nz = g.nz;
        dg.offset = voxels.size();
        size_t n = (size_t)g.nx * g.ny * g.nz;
        float mx = 0.0f;
        for (size_t k = 0; k < n; ++k) {
            voxels.push_back(g.density[k]);
            if (g.density[k] > mx) mx = g.density[k];
        }
        dg.max_density = mx;
        grids.push_back(dg);
    }
    std::vector<DevVolume> vols;
    bool tex_media = false;
    for (uint32_t i = 0; i < d->volume_count; ++i) {
        const ymrt_volume_desc &v = d->volumes[i];
        YM_REQUIRE(v.density_grid < 0 || (uint32_t)v.density_grid < d->grid_count, "volume grid out of range");
        YM_REQUIRE(v.density_texture < 0 || (uint32_t)v.density_texture < d->texture_count,
                   "volume density texture out of range");
        YM_REQUIRE(v.density_grid < 0 || v.density_texture < 0, "a volume takes a density grid or a density texture");
        DevVolume dv;
        dv.tex = v.density_texture;
        dv.tex_bound = v.density_texture >= 0 ? texture_upper_bound(d->textures[v.density_texture]) : 0.0f;
        tex_media = tex_media || v.density_texture >= 0;
        for (int k = 0; k < 3; ++k) {
            dv.sigma_t[k] = v.sigma_t[k];
            dv.albedo[k] = v.albedo[k];
        }
        dv.g = v.g;
        dv.grid = v.density_grid;
        dv.w2o = affine_from_colmajor(v.world_to_object);
        vols.push_back(dv);
    }
    // ------------------------------------------------ shape lights (LoadShapeLight)
    std::vector<DevLight> lights;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const ymrt_instance_desc &id = d->instances[i];
        const ymrt_material_desc &lm = d->materials[id.material];
        if (lm.type != YMRT_MATERIAL_LIGHT) continue;
        const float irr = (lm.light_color[0] + lm.light_color[1] + lm.light_color[2]) * lm.intensity * kPi;
        const DevInstance &in = h.instances[i];
        if (in.flags & kInstSphere) {
            // ellipsoid area estimate of LoadShapeLight with the instance's axis
            // scales (lengths of the object-to-world columns)
            const float sx = sphere_axis_scale(in.o2w, 0), sy = sphere_axis_scale(in.o2w, 1),
                        sz = sphere_axis_scale(in.o2w, 2);
            const float a = in.radius * sx, b = in.radius * sy, c = in.radius * sz;
            const float area = (4.0f / 3.0f) * kPi * (a * b + b * c + c * a);
            lights.push_back(DevLight{i, kEmpty, irr * area, area});
            continue;
        }
        for (uint32_t t = 0; t < in.tri_count; ++t) {
            const ShadeTri &st = h.shade[in.shade_base + t];
            f3 p0 = xf_position(in.o2w, f3{st.p[0], st.p[1], st.p[2]}, nullptr);
            f3 p1 = xf_position(in.o2w, f3{st.p[3], st.p[4], st.p[5]}, nullptr);
            f3 p2 = xf_position(in.o2w, f3{st.p[6], st.p[7], st.p[8]}, nullptr);
            float area = length(cross(p1 - p0, p2 - p0));
            lights.push_back(DevLight{i, t, irr * area, 0.5f * area});
        }
    }
    std::vector<float> cdf(lights.size());
    float total = 0.0f;
    for (size_t i = 0; i < lights.size(); ++i) {
        total += lights[i].power;
        cdf[i] = total;
    }
    float trcp = safe_rcp(total);
    for (size_t i = 0; i < cdf.size(); ++i) cdf[i] = total == 0.0f ? 0.0f : cdf[i] * trcp;
    std::vector<DevLightRec> light_recs(lights.size());
    std::vector<uint32_t> light_material(lights.size());
    for (size_t k = 0; k < lights.size(); ++k) {
        const DevLight &l = lights[k];
        const DevInstance &in = h.instances[l.instance];
        DevLightRec &r = light_recs[k];
        std::memset(&r, 0, sizeof(r));
        r.w2o = in.w2o;
        r.o2w = in.o2w;
        if (l.triangle != kEmpty) {
            const ShadeTri &st = h.shade[in.shade_base + l.triangle];
            std::memcpy(r.p, st.p, sizeof(r.p));
            std::memcpy(r.n, st.n, sizeof(r.n));
        }
        light_material[k] = in.material;
        light_emission_of(h.materials[in.material], r.le);
        r.area = l.area;
        r.radius = in.radius;
        r.triangle = l.triangle;
    }
    // ------------------------------------------------------------ distant lights
    std::vector<DevDistant> distant;
    YM_REQUIRE(d->distant_light_count == 0 || d->distant_lights != nullptr, "distant_lights is NULL");
    for (uint32_t i = 0; i < d->distant_light_count; ++i) {
        const ymrt_distant_light_desc &dl = d->distant_lights[i];
        DevDistant dd;
        dd.color = f3{dl.light_color[0], dl.light_color[1], dl.light_color[2]};
        dd.intensity = dl.intensity;
        // DistantLight() (src/Light.h:335-339): glm::cos(glm::radians(theta_max))
        dd.cos_max = cosf(dl.theta_max * 0.01745329251994329576923690768489f);
        dd.o2w = affine_from_colmajor(dl.object_to_world);
        dd.w2o = affine_from_colmajor(dl.world_to_object);
        distant.push_back(dd);
    }
    // ------------------------------------------------- Halton permutation table
    std::vector<uint32_t> primes, prime_sums(1001, 0);
    for (uint32_t c = 2; primes.size() < 1000; ++c) {
        bool p = true;
        for (uint32_t k = 2; k * k <= c; ++k)
            if (c % k == 0) {
                p = false;
                break;
            }
        if (p) primes.push_back(c);
    }
    for (int i = 0; i < 1000; ++i) prime_sums[i + 1] = prime_sums[i] + primes[i];
    std::vector<uint32_t> table(prime_sums[1000]);
    {
        std::mt19937 rng(d->halton_seed);
        for (int i = 0; i < 1000; ++i) {
            uint32_t p = primes[i];
            uint32_t *perm = table.data() + prime_sums[i];
            for (uint32_t j = 0; j < p; ++j) perm[j] = j;
            for (uint32_t j = p - 1; j > 0; --j) {
                uint32_t k = (uint32_t)rng() % (j + 1);
                std::swap(perm[j], perm[k]);
            }
        }
    }
    // ------------------------------------------------------------------ camera
    camera_setup(d, ds.camera);
    {
        DevScene hs = ds;
        hs.nodes = h.nodes.data();
        hs.tris = h.tris.data();
        hs.instances = h.instances.data();
        hs.shade = h.shade.data();
        camera_media(hs, h, d, ds.camera, sc->max_stack);
        sc->info.camera_medium_records = (uint32_t)ds.camera.record_count;
    }
    // ------------------------------------------------------------------ upload
    int rc;
    if ((rc = upload(sc, h.nodes, &ds.nodes))) return rc;
    if ((rc = upload(sc, h.tris, &ds.tris))) return rc;
    if ((rc = upload(sc, h.instances, &ds.instances))) return rc;
    if ((rc = upload(sc, h.shade, &ds.shade))) return rc;
    if ((rc = upload(sc, h.materials, &ds.materials))) return rc;
    {
        // the adjoint's per-material target slots, written by each adjoint call
        std::vector<int32_t> slots(2 * h.materials.size(), -1);
        const int32_t *d = nullptr;
        if ((rc = upload(sc, slots, &d))) return rc;
        sc->d_adj_slots = const_cast<int32_t *>(d);
        std::vector<uint32_t> dif(h.materials.size());
        for (size_t i = 0; i < h.materials.size(); ++i) dif[i] = shading_class(h.materials[i]);
        if ((rc = upload(sc, dif, &ds.mat_class))) return rc;
    }
    sc->materials = h.materials;
    sc->texture_count = d->texture_count;
    if ((rc = upload(sc, texs, &ds.textures))) return rc;
    sc->textures = texs;
    if ((rc = upload(sc, tex_prog, &ds.tex_prog))) return rc;
    if ((rc = upload(sc, texels, &ds.texels))) return rc;
    if ((rc = upload(sc, texels_f32, &ds.texels_f32))) return rc;
    if ((rc = upload(sc, lut, &ds.tex_lut))) return rc;
    if ((rc = upload(sc, grids, &ds.grids))) return rc;
    if ((rc = upload(sc, voxels, &ds.voxels))) return rc;
    if ((rc = upload(sc, vols, &ds.volumes))) return rc;
    sc->grids = grids;
    sc->volumes = vols;
    if ((rc = upload(sc, light_recs, &ds.light_recs))) return rc;
    if ((rc = upload(sc, cdf, &ds.light_cdf))) return rc;
    for (uint32_t i = 0; i < kLightCdfInline; ++i) ds.light_cdf_inline[i] = i < cdf.size() ? cdf[i] : 1.0f;
    sc->light_recs = light_recs;
    sc->light_material = light_material;
    if ((rc = upload(sc, table, &ds.halton))) return rc;
    if ((rc = upload(sc, distant, &ds.distant))) return rc;
    ds.distant_count = (uint32_t)distant.size();
    std::vector<HaltonDim> hdims(1000);
    for (int i = 0; i < 1000; ++i) hdims[i] = HaltonDim{primes[i], prime_sums[i], 1.0 / (double)primes[i]};
    if ((rc = upload(sc, hdims, &ds.halton_dims))) return rc;
    ds.texture_count = d->texture_count;
    ds.grid_count = d->grid_count;
    ds.has_tex_media = tex_media ? 1 : 0;
    ds.volume_count = d->volume_count;
    ds.light_count = (uint32_t)lights.size();
    sc->info.shape_light_count = ds.light_count;
    {
        void *p = nullptr;
        YM_HIP(hipMalloc(&p, 64));
        YM_HIP(hipMemset(p, 0, 64));
        sc->allocations.push_back(p);
        sc->d_error = static_cast<uint32_t *>(p);
    }
    YM_HIP(hipDeviceSynchronize());
    auto t1 = std::chrono::high_resolution_clock::now();
    sc->info.build_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
    sc->info.max_stack = sc->max_stack;
    return YMRT_OK;
}

// SceneModule::UpdateMaterial (src/core/SceneModule.cpp:654-663) for a range:
// the new records replace the old ones in HBM, and the flags derived from
// materials (nearby-hit list, texture-graph shading, Lambert-only) follow.
// Like the reference, the shape-light CDF and the camera's medium records
// keep what the scene was built with, so a material keeps its type here.
int update_materials(ymrt_scene *sc, uint32_t first, uint32_t count, const ymrt_material_desc *m, hipStream_t s)
{
    YM_REQUIRE(count == 0 || m != nullptr, "materials is NULL");
    YM_REQUIRE(first <= sc->materials.size() && count <= sc->materials.size() - first,
               "material range out of bounds");
    if (count == 0) return YMRT_OK;
    std::vector<DevMaterial> next = sc->materials;
    for (uint32_t i = 0; i < count; ++i) {
        DevMaterial dm;
        std::memcpy(&dm, &m[i], sizeof(dm));
        YM_REQUIRE(dm.type == next[first + i].type, "a material update cannot change the material type");
        YM_REQUIRE(material_textures_valid(dm, sc->texture_count), "material texture index out of range");
        next[first + i] = dm;
    }
    std::vector<uint32_t> dif(count);
    for (uint32_t i = 0; i < count; ++i) dif[i] = shading_class(next[first + i]);
    YM_HIP(hipMemcpyAsync(const_cast<DevMaterial *>(sc->dev.materials) + first, next.data() + first,
                          sizeof(DevMaterial) * count, hipMemcpyHostToDevice, s));
    YM_HIP(hipMemcpyAsync(const_cast<uint32_t *>(sc->dev.mat_class) + first, dif.data(), sizeof(uint32_t) * count,
                          hipMemcpyHostToDevice, s));
    // the shape lights of an updated light material emit its new radiance
    bool lights_changed = false;
    for (size_t k = 0; k < sc->light_recs.size(); ++k) {
        const uint32_t lm = sc->light_material[k];
        if (lm < first || lm >= first + count) continue;
        light_emission_of(next[lm], sc->light_recs[k].le);
        lights_changed = true;
    }
    if (lights_changed)
        YM_HIP(hipMemcpyAsync(const_cast<DevLightRec *>(sc->dev.light_recs), sc->light_recs.data(),
                              sizeof(DevLightRec) * sc->light_recs.size(), hipMemcpyHostToDevice, s));
    YM_HIP(hipStreamSynchronize(s));
    sc->materials.swap(next);
    sc->needs_nearby = needs_nearby(sc->has_boundaries, sc->materials);
    sc->dev.has_graph = uses_texture_graph(sc->procedural_textures, sc->materials) ? 1 : 0;
    return YMRT_OK;
}

int update_grid(ymrt_scene *sc, uint32_t grid, const float *density, hipStream_t s)
{
    YM_REQUIRE(density != nullptr, "density is NULL");
    YM_REQUIRE(grid < sc->grids.size(), "grid index out of range");
    DevGrid &g = sc->grids[grid];
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    float mx = 0.0f;
    for (size_t k = 0; k < n; ++k) mx = density[k] > mx ? density[k] : mx;
    g.max_density = mx;  // the majorant of the volumes that read the grid
    YM_HIP(hipMemcpyAsync(const_cast<float *>(sc->dev.voxels) + g.offset, density, sizeof(float) * n,
                          hipMemcpyHostToDevice, s));
    YM_HIP(hipMemcpyAsync(const_cast<DevGrid *>(sc->dev.grids) + grid, &g, sizeof(DevGrid), hipMemcpyHostToDevice, s));
    YM_HIP(hipStreamSynchronize(s));
    return YMRT_OK;
}

int update_texture(ymrt_scene *sc, uint32_t tex, const float *texels, hipStream_t s)
{
    YM_REQUIRE(texels != nullptr, "texels is NULL");
    YM_REQUIRE(tex < sc->textures.size(), "texture index out of range");
    const DevTexture &t = sc->textures[tex];
    YM_REQUIRE(t.type == YMRT_TEXTURE_IMAGE && t.f32, "only float image textures take texel updates");
    const size_t n = (size_t)t.width * t.height * t.channels;
    YM_HIP(hipMemcpyAsync(const_cast<float *>(sc->dev.texels_f32) + t.offset, texels, sizeof(float) * n,
                          hipMemcpyHostToDevice, s));
    YM_HIP(hipStreamSynchronize(s));
    return YMRT_OK;
}

}  // namespace ym

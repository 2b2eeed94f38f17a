// adjoint.cpp -- the targets of ymrt_render_adjoint: which parameters the
// path-replay adjoint can differentiate in a scene, and the replay pass's
// device-side description of them.
#include <cstring>
#include <string>

#include "internal.hpp"
#include "pt.hpp"

namespace ym {

// Lambert-only DefaultMtl: its one lobe is the diffuse albedo (specular,
// metal, transmission and coat weights 0, untextured), so cosine sampling and
// the MIS pdfs do not depend on the albedo and every later deposit of a path
// is linear in it.
static bool lambert_lobe(const DevMaterial &m)
{
    return m.type == YMRT_MATERIAL_DEFAULT && m.specular_weight == 0.0f && m.metalness == 0.0f &&
           m.transmission_weight == 0.0f && m.coat_weight == 0.0f && m.specular_weight_tex == kEmpty &&
           m.metalness_tex == kEmpty && m.transmission_weight_tex == kEmpty && m.coat_weight_tex == kEmpty;
}

static bool lambert_only(const DevMaterial &m) { return lambert_lobe(m) && m.diffuse_albedo_tex == kEmpty; }

// Metalness 1 without coat: the BSDF is the GGX conductor lobe alone; its
// roughness isotropic and untextured.
static bool metal_only(const DevMaterial &m)
{
    return m.type == YMRT_MATERIAL_DEFAULT && m.metalness == 1.0f && m.metalness_tex == kEmpty &&
           m.coat_weight == 0.0f && m.coat_weight_tex == kEmpty && m.roughness_x == m.roughness_y &&
           m.roughness_x_tex == kEmpty && m.roughness_y_tex == kEmpty;
}

// A material slot other than the diffuse albedo that reads texture t.
static bool other_use(const DevMaterial &m, uint32_t t)
{
    const uint32_t slots[] = {m.specular_albedo_tex, m.roughness_x_tex,   m.roughness_y_tex,   m.specular_weight_tex,
                              m.metalness_tex,       m.transmission_weight_tex, m.coat_albedo_tex, m.coat_weight_tex,
                              m.coat_thickness_tex,  m.coat_roughness_x_tex, m.coat_roughness_y_tex,
                              m.normal_mapping_tex,  m.bump_mapping_tex};
    for (uint32_t s : slots)
        if (s == t) return true;
    return false;
}

int prepare_adjoint(ymrt_scene *sc, const ymrt_adjoint_targets *tg, const float *adj_img, hipStream_t stream,
                    AdjParams *out)
{
    YM_REQUIRE(tg != nullptr, "NULL adjoint targets");
    YM_REQUIRE(adj_img != nullptr, "NULL adjoint image");
    const uint32_t nm = (uint32_t)sc->materials.size();
    std::vector<int32_t> slots(2 * (size_t)nm, -1);
    YM_REQUIRE(tg->albedo_count == 0 || tg->albedo_materials != nullptr, "NULL albedo_materials");
    YM_REQUIRE(tg->albedo_count == 0 || tg->grad_albedo != nullptr, "NULL grad_albedo");
    for (uint32_t k = 0; k < tg->albedo_count; ++k) {
        const uint32_t m = tg->albedo_materials[k];
        YM_REQUIRE(m < nm, "albedo target material " + std::to_string(m) + " out of range");
        YM_REQUIRE(lambert_only(sc->materials[m]),
                   "the adjoint cannot differentiate the albedo of material " + std::to_string(m) +
                       ": not a Lambert-only DefaultMtl with an untextured albedo");
        YM_REQUIRE(slots[m] < 0, "albedo target material " + std::to_string(m) + " listed twice");
        slots[m] = (int32_t)k;
    }
    YM_REQUIRE(tg->roughness_count == 0 || tg->roughness_materials != nullptr, "NULL roughness_materials");
    YM_REQUIRE(tg->roughness_count == 0 || tg->grad_roughness != nullptr, "NULL grad_roughness");
    for (uint32_t k = 0; k < tg->roughness_count; ++k) {
        const uint32_t m = tg->roughness_materials[k];
        YM_REQUIRE(m < nm, "roughness target material " + std::to_string(m) + " out of range");
        YM_REQUIRE(metal_only(sc->materials[m]),
                   "the adjoint cannot differentiate the roughness of material " + std::to_string(m) +
                       ": not a metal-only DefaultMtl with an isotropic untextured roughness");
        YM_REQUIRE(slots[nm + m] < 0, "roughness target material " + std::to_string(m) + " listed twice");
        slots[nm + m] = (int32_t)k;
    }
    const uint32_t tex = tg->texture;
    if (tex != YMRT_INVALID_INDEX) {
        YM_REQUIRE(tex < sc->textures.size(), "texture target " + std::to_string(tex) + " out of range");
        YM_REQUIRE(sc->textures[tex].type == YMRT_TEXTURE_IMAGE,
                   "the adjoint cannot differentiate texture " + std::to_string(tex) + ": not an image texture");
        YM_REQUIRE(tg->grad_texture != nullptr, "NULL grad_texture");
        for (uint32_t m = 0; m < nm; ++m) {
            const DevMaterial &md = sc->materials[m];
            YM_REQUIRE(!other_use(md, tex) && (md.diffuse_albedo_tex != tex || lambert_lobe(md)),
                       "the adjoint cannot differentiate texture " + std::to_string(tex) + ": material " +
                           std::to_string(m) + " uses it other than as the albedo of a Lambert-only DefaultMtl");
        }
        for (const DevVolume &v : sc->volumes)
            YM_REQUIRE(v.tex != (int32_t)tex, "the adjoint cannot differentiate texture " + std::to_string(tex) +
                                                  ": a volume's density reads it");
    }
    const uint32_t vol = tg->volume;
    if (vol != YMRT_INVALID_INDEX) {
        YM_REQUIRE(vol < sc->volumes.size(), "volume target " + std::to_string(vol) + " out of range");
        const DevVolume &v = sc->volumes[vol];
        YM_REQUIRE(v.grid >= 0 && (size_t)v.grid < sc->grids.size() && v.tex < 0,
                   "the adjoint cannot differentiate volume " + std::to_string(vol) + ": no voxel-grid density");
        YM_REQUIRE(v.sigma_t[0] == v.sigma_t[1] && v.sigma_t[0] == v.sigma_t[2],
                   "the adjoint cannot differentiate volume " + std::to_string(vol) + ": sigma_t is not gray");
        YM_REQUIRE(tg->grad_density != nullptr, "NULL grad_density");
    }
    if (nm) {
        YM_HIP(hipMemcpyAsync(sc->d_adj_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice,
                              stream));
        YM_HIP(hipStreamSynchronize(stream));
    }
    std::memset(out, 0, sizeof(*out));
    out->img = adj_img;
    out->alb_slot = sc->d_adj_slots;
    out->rough_slot = sc->d_adj_slots + nm;
    out->g_alb = tg->grad_albedo;
    out->g_rough = tg->grad_roughness;
    out->g_tex = tg->grad_texture;
    out->g_vox = tg->grad_density;
    out->n_alb = tg->albedo_count;
    out->n_rough = tg->roughness_count;
    out->tex = tex;
    out->vol = vol;
    out->lds = (size_t)tg->albedo_count * 3 + tg->roughness_count <= kAdjLds ? 1u : 0u;
    return YMRT_OK;
}

}  // namespace ym

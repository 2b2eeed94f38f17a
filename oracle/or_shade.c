/*
 * or_shade.c -- oracle restatement of YMRT's shading and path loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path tracer; never shipped).
 *
 *   - Camera basis / GenerateRay: src/Camera.h:164-170, 311-322
 *   - InitCameraRayTransfer: src/Camera.cpp:15-78
 *   - ray generation (pixel/sample -> Halton or PCG sampler, jitter):
 *     src/Shading.cu:1634-1725
 *   - trace stage with volume free flight, miss -> background:
 *     src/Shading.cu:1765-1901, Background 35-38
 *   - shading stage: src/Shading.cu:1953-2302
 *   - NEE: EvalShapeLight 294-417, EvalVolumeShapeLight 419-513, PowerHeuristic
 *     40-43, indirect_light_clamp 61-81
 *   - FetchMeshShadingData: src/ShadingUtilities.cuh:713-875
 *   - BoundaryTransitionBunch: src/ShadingUtilities.cuh:954-990, RayTransfer
 *     src/Ray.h:106-260
 *   - BSDF: src/BSDF.cuh:20-899 (lobes, DefaultMtlBSDF, MaterialBSDF without
 *     normal or bump maps)
 *   - lights: src/Light.h:14-40 (RayIntersectTriangle), 87-321 (ShapeLight
 *     mesh branch), 398-416 (SampleLightPower); CDF
 *     src/core/SceneModule.cpp:1203-1304
 *   - sampling helpers: src/SampleUtilities.h:11-48
 *   - volumes: src/VolumeEval.cuh:18-538, src/Volume.h:21-48
 *   - image textures: src/TextureEval.cuh:162-237, decode src/FileIO.cpp:17-45,
 *     src/MathCommon.h:136-139
 *   - accumulate / tonemap: src/Shading.cu:2354-2425
 *
 * Two documented departures, shared with the HIP path:
 *   - heterogeneous media use a voxel grid (the reference has none); the
 *     shadow-ray transmittance through such a grid draws its random numbers
 *     from a PCG hash of (pixel, sample, frame, vertex, ray, step) instead of
 *     the path's Halton stream;
 *   - the shape-light list follows input instance order (the reference's
 *     follows its BVH-reordered instance order).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle.h"

/* ------------------------------------------------------------------ setup */
float or_srgb_to_linear(float s)
{
    return s > 0.04045 ? (float)pow(((double)s + 0.055) / 1.055, (double)2.4f) : (s / 12.92f);
}

static void camera_basis(or_scene *sc)
{
    v3 w = mk3(sc->camera.w[0], sc->camera.w[1], sc->camera.w[2]);
    w = scl3(w, safe_rcp(length3(w)));
    v3 up = mk3(0.0f, 1.0f, 0.0f);
    double mu[3], mv[3];
    mu[0] = (double)up.y * (double)w.z - (double)up.z * (double)w.y;
    mu[1] = -((double)up.x * (double)w.z - (double)up.z * (double)w.x);
    mu[2] = (double)up.x * (double)w.y - (double)up.y * (double)w.x;
    v3 u = scl3(mk3((float)mu[0], (float)mu[1], (float)mu[2]),
                safe_rcp((float)sqrt(mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2])));
    mv[0] = (double)w.y * (double)u.z - (double)w.z * (double)u.y;
    mv[1] = -((double)w.x * (double)u.z - (double)w.z * (double)u.x);
    mv[2] = (double)w.x * (double)u.y - (double)w.y * (double)u.x;
    v3 v = scl3(mk3((float)mv[0], (float)mv[1], (float)mv[2]),
                safe_rcp((float)sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2])));
    sc->cam_u = u; sc->cam_v = v; sc->cam_w = w;
    sc->tan_half_fov = tanf(0.5f * sc->camera.fov);
}

static or_ray make_ray(v3 o, v3 d, uint32_t li, uint32_t lt)
{
    or_ray r;
    r.o = o; r.d = normalize3(d); r.t = OR_TMAX; r.last_inst = li; r.last_tri = lt;
    return r;
}

static or_ray camera_ray(const or_scene *sc, float nx, float ny)
{
    float height = 2.0f * sc->tan_half_fov;
    float width = height * sc->camera.aspect_ratio;
    v3 dir = add3(add3(neg3(sc->cam_w), sclf3((nx - 0.5f) * width, sc->cam_u)),
                  sclf3((ny - 0.5f) * height, sc->cam_v));
    return make_ray(mk3(sc->camera.position[0], sc->camera.position[1], sc->camera.position[2]), dir,
                    OR_EMPTY, OR_EMPTY);
}

/* --------------------------------------------------------------- transfer */
static float fetch_ior(const ymrt_material_desc *m, uint32_t *prio)
{
    if (m->type == YMRT_MATERIAL_DEFAULT) { *prio = m->ior_priority; return m->ior_n; }
    *prio = 0;
    return 1.0f;
}

static int tr_find_max(const or_transfer *t)
{
    if (t->count == 0) return -1;
    uint32_t mp = t->rec[0].priority;
    int mi = 0;
    for (int i = 1; i < t->count; ++i)
        if (t->rec[i].priority >= mp) { mp = t->rec[i].priority; mi = i; }
    return mi;
}
static int tr_find_prim(const or_transfer *t, uint32_t prim)
{
    for (int i = t->count - 1; i >= 0; --i) if (t->rec[i].prim == prim) return i;
    return -1;
}
static int tr_push(or_transfer *t, uint32_t prim, float ior, uint32_t prio, int vol)
{
    if (t->count == OR_MAX_BOUNDARY_RECORD) return 0;
    or_record *r = &t->rec[t->count];
    r->prim = prim; r->ior = ior; r->priority = prio; r->vol = vol;
    ++t->count;
    if (t->count == 1) { t->max_idx = 0; return 1; }
    if (prio >= t->rec[t->max_idx].priority) t->max_idx = t->count - 1;
    return 1;
}
static int tr_pop(or_transfer *t, int idx)
{
    if (t->count == 0) return 0;
    if (!(idx >= 0 && idx < t->count)) return 0;
    for (int i = idx + 1; i < t->count; ++i) t->rec[i - 1] = t->rec[i];
    --t->count;
    t->max_idx = tr_find_max(t);
    return 1;
}
static int tr_volumes(const or_transfer *t, int *vols, uint32_t volume_count)
{
    if (t->count == 0 || t->max_idx == -1) return 0;
    uint32_t mp = t->rec[t->max_idx].priority;
    int n = 0;
    for (int i = t->count - 1; i >= 0; --i)
        if (t->rec[i].priority == mp && t->rec[i].vol >= 0 && (uint32_t)t->rec[i].vol < volume_count)
            vols[n++] = t->rec[i].vol;
    return n;
}
static float tr_ray_ior(const or_transfer *t) { return t->count > 0 ? t->rec[t->max_idx].ior : 1.0f; }
static float tr_ex_ior(const or_transfer *t, uint32_t hit_prim)
{
    uint32_t mp = 0;
    int mi = -1;
    for (int i = 0; i < t->count; ++i) {
        if (t->rec[i].prim == hit_prim) continue;
        if (t->rec[i].priority >= mp) { mp = t->rec[i].priority; mi = i; }
    }
    return mi == -1 ? 1.0f : t->rec[mi].ior;
}

static void boundary_bunch(const or_scene *sc, or_transfer *t, const or_nearby *nb, int n)
{
    for (int i = 0; i < n; ++i) {
        if (nb[i].back) continue;
        const ymrt_instance_desc *in = &sc->instances[nb[i].prim];
        uint32_t prio;
        float ior = fetch_ior(&sc->materials[in->material], &prio);
        if (tr_find_prim(t, nb[i].prim) == -1) tr_push(t, nb[i].prim, ior, prio, in->inner_volume);
    }
    for (int i = 0; i < n; ++i) {
        if (!nb[i].back) continue;
        int ri = tr_find_prim(t, nb[i].prim);
        if (ri != -1) tr_pop(t, ri);
    }
}

/* src/Camera.cpp:15-78.  The reference walks an unordered_map; here the
 * distinct instances are visited in the order the centre ray meets them. */
static void camera_transfer(or_scene *sc)
{
    or_transfer *ct = &sc->camera_transfer;
    ct->count = 0; ct->max_idx = -1;
    uint32_t seen[48];
    int ns = 0;
    or_ray ray = camera_ray(sc, 0.5f, 0.5f);
    for (int depth = 0; depth < 48; ++depth) {
        or_hit h;
        memset(&h, 0, sizeof(h));
        h.t = OR_TMAX; h.inst = OR_EMPTY; h.tri = OR_EMPTY;
        or_nearby nb[OR_MAX_NEARBY];
        int nc;
        if (!or_trace_ray(sc, &ray, &h, &nc, nb)) break;
        v3 p, perr, ng;
        or_fetch_geometry_normal(sc, &h, &p, &perr, &ng);
        int found = 0;
        for (int i = 0; i < ns; ++i) if (seen[i] == h.inst) found = 1;
        if (!found) seen[ns++] = h.inst;
        ray = make_ray(or_offset_ray_origin(p, perr, ray.d, ng), ray.d, h.inst, h.tri);
    }
    for (int i = 0; i < ns; ++i) {
        const ymrt_instance_desc *in = &sc->instances[seen[i]];
        or_ray test = camera_ray(sc, 0.5f, 0.5f);
        test.last_inst = OR_EMPTY; test.last_tri = OR_EMPTY;
        or_ray pos, negr;
        pos.o = m4_point(in->world_to_object, test.o);
        pos.d = m4_dir(in->world_to_object, test.d);
        pos.t = test.t; pos.last_inst = OR_EMPTY; pos.last_tri = OR_EMPTY;
        negr = pos;
        negr.d = neg3(pos.d);
        /* any-hit against this instance only */
        int hp = 0, hn = 0;
        if (in->geometry_type == YMRT_GEOMETRY_SPHERE) {
            hp = or_sphere_any(sc->spheres[in->geometry].radius, &pos);
            hn = or_sphere_any(sc->spheres[in->geometry].radius, &negr);
        }
        const or_mesh *m = in->geometry_type == YMRT_GEOMETRY_SPHERE ? NULL : &sc->meshes[in->geometry];
        for (uint32_t t = 0; m && t < m->tri_count && !(hp && hn); ++t) {
            v3 p0 = mk3(m->pos[3 * m->pidx[3 * t]], m->pos[3 * m->pidx[3 * t] + 1], m->pos[3 * m->pidx[3 * t] + 2]);
            v3 p1 = mk3(m->pos[3 * m->pidx[3 * t + 1]], m->pos[3 * m->pidx[3 * t + 1] + 1], m->pos[3 * m->pidx[3 * t + 1] + 2]);
            v3 p2 = mk3(m->pos[3 * m->pidx[3 * t + 2]], m->pos[3 * m->pidx[3 * t + 2] + 1], m->pos[3 * m->pidx[3 * t + 2] + 2]);
            float th; v3 b; int bk;
            if (!hp && or_intersect_triangle(p0, p1, p2, &pos, &th, &b, &bk) && th >= OR_TMIN && th <= pos.t) hp = 1;
            if (!hn && or_intersect_triangle(p0, p1, p2, &negr, &th, &b, &bk) && th >= OR_TMIN && th <= negr.t) hn = 1;
        }
        if (hp && hn) {
            uint32_t prio;
            float ior = fetch_ior(&sc->materials[in->material], &prio);
            if (!tr_push(ct, seen[i], ior, prio, in->inner_volume)) break;
        }
    }
}

static v3 ld3(const float *a, uint32_t i) { return mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

void or_shading_init(or_scene *sc, const ymrt_scene_desc *d)
{
    sc->material_count = d->material_count;
    sc->materials = d->materials;
    sc->volume_count = d->volume_count;
    sc->volumes = d->volumes;
    sc->camera = d->camera;
    /* textures */
    sc->texture_count = d->texture_count;
    sc->textures = (or_texture *)calloc(d->texture_count ? d->texture_count : 1, sizeof(or_texture));
    float lut[2][256];
    for (int i = 0; i < 256; ++i) {
        float v = (float)i * (1.0f / 255.0f);
        lut[0][i] = or_srgb_to_linear(v);
        lut[1][i] = v;
    }
    for (uint32_t ti = 0; ti < d->texture_count; ++ti) {
        const ymrt_texture_desc *td = &d->textures[ti];
        or_texture *t = &sc->textures[ti];
        t->desc = td;
        if (td->type != YMRT_TEXTURE_IMAGE) continue;
        t->width = td->width; t->height = td->height; t->channels = td->channels;
        t->wrap = td->wrap_mode; t->color_space = td->color_space;
        t->su = td->scale_u; t->sv = td->scale_v; t->ou = td->offset_u; t->ov = td->offset_v;
        size_t n = (size_t)td->width * td->height * td->channels;
        t->texels = (float *)malloc(sizeof(float) * (n ? n : 1));
        if (td->texels_f32) {
            /* linear float texels are taken as they are (no colour-space
             * conversion) */
            for (size_t i = 0; i < n; ++i) t->texels[i] = td->texels_f32[i];
        } else {
            const float *l = lut[td->color_space == YMRT_COLOR_SRGB ? 0 : 1];
            for (size_t i = 0; i < n; ++i) t->texels[i] = l[td->texels[i]];
        }
    }
    sc->grid_count = d->grid_count;
    sc->grids = (or_grid *)calloc(d->grid_count ? d->grid_count : 1, sizeof(or_grid));
    for (uint32_t gi = 0; gi < d->grid_count; ++gi) {
        or_grid *g = &sc->grids[gi];
        g->nx = d->grids[gi].nx; g->ny = d->grids[gi].ny; g->nz = d->grids[gi].nz;
        g->density = d->grids[gi].density;
        float mx = 0.0f;
        size_t n = (size_t)g->nx * g->ny * g->nz;
        for (size_t i = 0; i < n; ++i) if (g->density[i] > mx) mx = g->density[i];
        g->max_density = mx;
    }
    /* distant lights: DistantLight(), src/Light.h:335-339 */
    sc->distant_count = d->distant_light_count;
    sc->distant = d->distant_lights;
    sc->distant_cos_max = (float *)calloc(d->distant_light_count ? d->distant_light_count : 1, sizeof(float));
    for (uint32_t i = 0; i < d->distant_light_count; ++i)
        sc->distant_cos_max[i] = cosf(d->distant_lights[i].theta_max * 0.01745329251994329576923690768489f);
    /* shape lights: src/core/SceneModule.cpp:1203-1304 */
    uint32_t nl = 0;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const ymrt_instance_desc *in = &d->instances[i];
        if (in->material < d->material_count && d->materials[in->material].type == YMRT_MATERIAL_LIGHT)
            nl += in->geometry_type == YMRT_GEOMETRY_SPHERE ? 1 : d->meshes[in->geometry].triangle_count;
    }
    sc->light_count = nl;
    sc->lights = (or_shape_light *)calloc(nl ? nl : 1, sizeof(or_shape_light));
    sc->light_cdf = (float *)calloc(nl ? nl : 1, sizeof(float));
    nl = 0;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const ymrt_instance_desc *in = &d->instances[i];
        if (!(in->material < d->material_count && d->materials[in->material].type == YMRT_MATERIAL_LIGHT)) continue;
        const ymrt_material_desc *lm = &d->materials[in->material];
        float irr = (lm->light_color[0] + lm->light_color[1] + lm->light_color[2]) * lm->intensity * OR_PI;
        if (in->geometry_type == YMRT_GEOMETRY_SPHERE) {
            /* sphere branch: ellipsoid area estimate with the instance's axis
             * scales (lengths of the object-to-world columns) */
            const float *m = in->object_to_world;
            float sx = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            float sy = sqrtf(m[4] * m[4] + m[5] * m[5] + m[6] * m[6]);
            float sz = sqrtf(m[8] * m[8] + m[9] * m[9] + m[10] * m[10]);
            float radius = d->spheres[in->geometry].radius;
            float a = radius * sx, b = radius * sy, c = radius * sz;
            float area = (4.0f / 3.0f) * OR_PI * (a * b + b * c + c * a);
            sc->lights[nl].instance = i;
            sc->lights[nl].triangle = OR_EMPTY;
            sc->lights[nl].power = irr * area;
            sc->lights[nl].area = area;
            ++nl;
            continue;
        }
        const ymrt_mesh_desc *md = &d->meshes[in->geometry];
        for (uint32_t t = 0; t < md->triangle_count; ++t) {
            v3 p0 = transform_position(in->object_to_world, ld3(md->positions, md->position_indices[3 * t]), NULL);
            v3 p1 = transform_position(in->object_to_world, ld3(md->positions, md->position_indices[3 * t + 1]), NULL);
            v3 p2 = transform_position(in->object_to_world, ld3(md->positions, md->position_indices[3 * t + 2]), NULL);
            float area = length3(cross3(sub3(p1, p0), sub3(p2, p0)));
            sc->lights[nl].instance = i;
            sc->lights[nl].triangle = t;
            sc->lights[nl].power = irr * area;
            sc->lights[nl].area = 0.5f * area;
            ++nl;
        }
    }
    float total = 0.0f;
    for (uint32_t i = 0; i < nl; ++i) { total += sc->lights[i].power; sc->light_cdf[i] = total; }
    float trcp = safe_rcp(total);
    for (uint32_t i = 0; i < nl; ++i) sc->light_cdf[i] = total == 0.0f ? 0.0f : sc->light_cdf[i] * trcp;
    /* sampler table */
    sc->halton = (uint32_t *)malloc(sizeof(uint32_t) * or_halton_table_size());
    or_halton_table(d->halton_seed, sc->halton);
    camera_basis(sc);
    camera_transfer(sc);
}

void or_shading_free(or_scene *sc)
{
    free(sc->distant_cos_max);
    for (uint32_t i = 0; i < sc->texture_count; ++i) free(sc->textures[i].texels);
    free(sc->textures);
    free(sc->grids);
    free(sc->lights);
    free(sc->light_cdf);
    free(sc->halton);
}

/* ---------------------------------------------------------------- textures */
static v3 texel(const or_texture *t, int x, int y)
{
    int w = (int)t->width, h = (int)t->height;
    if (t->wrap == YMRT_WRAP_CLAMP) {
        x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
        y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    } else {
        x = ((x % w) + w) % w;
        y = ((y % h) + h) % h;
    }
    const float *p = t->texels + ((size_t)y * w + x) * t->channels;
    if (t->channels == 1) return s3(p[0]);
    v3 r = s3(0.0f);
    for (uint32_t c = 0; c < 3 && c < t->channels; ++c) setc3(&r, (int)c, p[c]);
    return r;
}

v3 or_texture_eval(const or_scene *sc, uint32_t ti, float u, float v)
{
    const or_texture *t = &sc->textures[ti];
    float tu = u * t->su + t->ou;
    if (tu < 0.0f) tu = fmaxr(0.0f, tu + fabsf(floorf(tu)));
    float tv = v * t->sv + t->ov;
    if (tv < 0.0f) tv = fmaxr(0.0f, tv + fabsf(floorf(tv)));
    float px = tu * (float)t->width, py = tv * (float)t->height;
    float cx = px - 0.5f, cy = py - 0.5f;
    int x0 = (int)floorf(cx), y0 = (int)floorf(cy);
    float fx = cx - (float)x0, fy = cy - (float)y0;
    int x1 = x0 + 1, y1 = y0 + 1;
    v3 r = sclf3((1.0f - fx) * (1.0f - fy), texel(t, x0, y0));
    r = add3(r, sclf3(fx * (1.0f - fy), texel(t, x1, y0)));
    r = add3(r, sclf3((1.0f - fx) * fy, texel(t, x0, y1)));
    r = add3(r, sclf3(fx * fy, texel(t, x1, y1)));
    return r;
}

/* -------------------------------------------------------------------- BSDF */
enum { LOBE_LAMBERT = 0, LOBE_REFLECT = 1, LOBE_TRANSMIT = 2 };
typedef struct {
    int type;
    v3 albedo;          /* diffuse / specular / transmission albedo */
    v3 nt, kt;          /* reflection */
    float ni, nts;      /* ni; transmission nt */
    float ax, ay;
    int metal;
} lobe;
typedef struct {
    v3 t, b, n;
    int is_default;
    lobe l[6];
    float w[6];
    int count;
} mbsdf;

static int same_hemi(v3 a, v3 b) { return a.z * b.z > 0.0f; }
static v3 reflect3(v3 wo, v3 n) { return add3(neg3(wo), scl3(sclf3(2.0f, n), dot3(n, wo))); }
static int refract3(v3 wo, v3 n, float ni, float nt, v3 *wi)
{
    float ci = fmaxr(dot3(wo, n), 0.0f);
    float ci2 = sqrf(ci);
    float si2 = 1.0f - ci2;
    float ct2 = 1.0f - sqrf(ni / nt) * si2;
    if (!(ct2 > 0.0f)) return 0;
    float ct = sqrtf(ct2);
    *wi = normalize3(sub3(sclf3(-(ni / nt), sub3(wo, scl3(n, ci))), scl3(n, ct)));
    return 1;
}
static float fresnel_dd(float eta, float c)
{
    c = fmaxr(0.0f, c);
    float s2 = 1.f - c * c;
    float t0 = sqrtf(fmaxr(1.0f - (s2 / (eta * eta)), 0.0f));
    float t1 = eta * t0;
    float t2 = eta * c;
    float rs = (c - t1) / fmaxr(c + t1, 1E-12f);
    float rp = (t0 - t2) / fmaxr(t0 + t2, 1E-12f);
    return (float)(0.5 * (double)(rs * rs + rp * rp));
}
static v3 fresnel_dc(v3 eta, v3 etak, float c)
{
    float c2 = c * c, s2 = 1.0f - c2;
    v3 e2 = mul3(eta, eta), k2 = mul3(etak, etak);
    v3 t0 = sub3(sub3(e2, k2), s3(s2));
    v3 ab = mk3(sqrtf(t0.x * t0.x + 4.0f * e2.x * k2.x), sqrtf(t0.y * t0.y + 4.0f * e2.y * k2.y),
                sqrtf(t0.z * t0.z + 4.0f * e2.z * k2.z));
    v3 t1 = add3(ab, s3(c2));
    v3 a = mk3(sqrtf(0.5f * (ab.x + t0.x)), sqrtf(0.5f * (ab.y + t0.y)), sqrtf(0.5f * (ab.z + t0.z)));
    v3 t2 = scl3(sclf3(2.0f, a), c);
    v3 rs = div3(sub3(t1, t2), add3(t1, t2));
    v3 t3 = add3(sclf3(c2, ab), s3(s2 * s2));
    v3 t4 = scl3(t2, s2);
    v3 rp = div3(mul3(rs, sub3(t3, t4)), add3(t3, t4));
    return sclf3(0.5f, add3(rp, rs));
}
static void edge_tint(v3 refl, v3 tint, v3 *n, v3 *k)
{
    v3 r = mk3(clampf(refl.x, 0.f, 0.99f), clampf(refl.y, 0.f, 0.99f), clampf(refl.z, 0.f, 0.99f));
    v3 rs = mk3(sqrtf(r.x), sqrtf(r.y), sqrtf(r.z));
    v3 one = s3(1.0f);
    v3 nmin = div3(sub3(one, r), add3(one, r));
    v3 nmax = div3(add3(one, rs), sub3(one, rs));
    *n = mixv3(nmax, nmin, tint);
    v3 np1 = add3(*n, one), nm1 = sub3(*n, one);
    v3 k2 = div3(sub3(mul3(mul3(np1, np1), r), mul3(nm1, nm1)), sub3(one, r));
    k2 = max3(k2, s3(0.0f));
    *k = mk3(sqrtf(k2.x), sqrtf(k2.y), sqrtf(k2.z));
}
static float ggx_d(v3 h, float ax, float ay)
{
    float den = OR_PI * ax * ay * sqrf(sqrf(h.x / ax) + sqrf(h.y / ay) + sqrf(h.z));
    return 1.0f / fmaxr(den, 1E-12f);
}
static float ggx_lambda(v3 w, float ax, float ay)
{
    if (fabsf(w.z) < 1E-8f) return 0.0f;
    return (-1.0f + sqrtf(1.0f + (sqrf(w.x * ax) + sqrf(w.y * ay)) / sqrf(w.z))) * 0.5f;
}
static float ggx_g1(v3 w, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(w, ax, ay)); }
static float ggx_g2(v3 a, v3 b, float ax, float ay)
{
    return 1.0f / (1.0f + ggx_lambda(a, ax, ay) + ggx_lambda(b, ax, ay));
}
static v3 ggx_sample_wh(v3 wo, float u0, float u1, float ax, float ay)
{
    v3 wh = normalize3(mk3(ax * wo.x, ay * wo.y, wo.z));
    float lensq = wh.x * wh.x + wh.y * wh.y;
    v3 T1 = lensq > 0 ? scl3(mk3(-wh.y, wh.x, 0), 1.0f / sqrtf(lensq)) : mk3(1, 0, 0);
    v3 T2 = cross3(wh, T1);
    float r = sqrtf(fmaxr(u0, 0.0f));
    float phi = 2.0f * OR_PI * u1;
    float t1 = r * or_cosf(phi);
    float t2 = r * or_sinf(phi);
    float s = (float)(0.5 * (1.0 + (double)wh.z));
    t2 = (float)((1.0 - (double)s) * sqrt(1.0 - (double)(t1 * t1)) + (double)s * (double)t2);
    v3 Nh = add3(add3(sclf3(t1, T1), sclf3(t2, T2)), sclf3(sqrtf(fmaxr(0.0f, 1.0f - t1 * t1 - t2 * t2)), wh));
    return normalize3(mk3(ax * Nh.x, ay * Nh.y, fmaxr(0.0f, Nh.z)));
}
static v3 sample_cosine(float u0, float u1)
{
    float r = sqrtf(u0);
    float phi = u1 * OR_TWO_PI;
    float x = r * or_cosf(phi), y = r * or_sinf(phi);
    float r2 = x * x + y * y;
    return mk3(x, y, r2 < 1.0f ? sqrtf(1.0f - r2) : 0.0f);
}

static v3 lobe_eval(const lobe *l, v3 wo, v3 wi, float *pdf)
{
    if (l->type == LOBE_LAMBERT) {
        if (!same_hemi(wo, wi)) { *pdf = 0.0f; return s3(0.0f); }
        float c = fmaxr(wi.z, 0.0f);
        *pdf = c * OR_INV_PI;
        return scl3(scl3(l->albedo, c), OR_INV_PI);
    }
    if (l->type == LOBE_REFLECT) {
        if (!same_hemi(wo, wi) || wo.z < 0.f || wi.z < 0.f) { *pdf = 0.0f; return s3(0.0f); }
        v3 wh = add3(wi, wo);
        float len = length3(wh);
        if (fabsf(len) < 1E-8f) { *pdf = 0.0f; return s3(0.0f); }
        wh = divs3(wh, len);
        float D = ggx_d(wh, l->ax, l->ay);
        float G1 = ggx_g1(wo, l->ax, l->ay);
        float G2 = ggx_g2(wo, wi, l->ax, l->ay);
        float ini = safe_rcp(l->ni);
        float c = fmaxr(dot3(wo, wh), 0.0f);
        v3 fr = l->metal ? fresnel_dc(scl3(l->nt, ini), scl3(l->kt, ini), c) : s3(fresnel_dd(l->nt.x * ini, c));
        *pdf = D * G1 * 0.25f / fmaxr(wo.z, 1E-8f);
        return divs3(mul3(scl3(scl3(scl3(l->albedo, D), G2), 0.25f), fr), fmaxr(wo.z, 1E-8f));
    }
    /* transmission */
    if (same_hemi(wo, wi)) { *pdf = 0.0f; return s3(0.0f); }
    v3 wh = neg3(add3(sclf3(l->ni, wo), sclf3(l->nts, wi)));
    if (!same_hemi(wo, wh)) { *pdf = 0.0f; return s3(0.0f); }
    float len = length3(wh);
    if (fabsf(len) < 1E-8f) { *pdf = 0.0f; return s3(0.0f); }
    wh = divs3(wh, len);
    float D = ggx_d(wh, l->ax, l->ay);
    float G1 = ggx_g1(wo, l->ax, l->ay);
    float G2 = ggx_g2(wo, wi, l->ax, l->ay);
    float dwh = sqrf(l->nts) * fabsf(dot3(wi, wh)) /
                fmaxr(sqrf(l->ni * dot3(wo, wh) + l->nts * dot3(wi, wh)), 1E-12f);
    float ini = safe_rcp(l->ni);
    float c = fmaxr(dot3(wo, wh), 0.0f);
    float fr = fresnel_dd(l->nts * ini, c);
    *pdf = D * G1 * fabsf(dot3(wo, wh)) * dwh / fmaxr(fabsf(wo.z), 1E-8f);
    v3 r = scl3(l->albedo, 1.0f - fr);
    r = scl3(r, fabsf(dot3(wo, wh) / fmaxr(fabsf(wo.z), 1E-8f)));
    r = scl3(r, G2); r = scl3(r, D); r = scl3(r, dwh); r = scl3(r, sqrf(l->ni / l->nts));
    return r;
}

static int lobe_sample(const lobe *l, float u0, float u1, v3 wo, v3 *weight, v3 *wi, float *pdf)
{
    if (l->type == LOBE_LAMBERT) {
        *wi = sample_cosine(u0, u1);
        float c = fmaxr(wi->z, 0.0f);
        *pdf = c * OR_INV_PI;
        *weight = l->albedo;
        return 1;
    }
    v3 wh = ggx_sample_wh(wo, u0, u1, l->ax, l->ay);
    if (!same_hemi(wo, wh) || dot3(wo, wh) < 0.0f) return 0;
    if (l->type == LOBE_REFLECT) {
        *wi = reflect3(wo, wh);
        if (!same_hemi(*wi, wo)) return 0;
        float D = ggx_d(wh, l->ax, l->ay);
        float G1 = ggx_g1(wo, l->ax, l->ay);
        float G2 = ggx_g2(wo, *wi, l->ax, l->ay);
        float ini = safe_rcp(l->ni);
        float c = fmaxr(dot3(wo, wh), 0.0f);
        v3 fr = l->metal ? fresnel_dc(scl3(l->nt, ini), scl3(l->kt, ini), c) : s3(fresnel_dd(l->nt.x * ini, c));
        *pdf = D * G1 * 0.25f / fmaxr(wo.z, 1E-8f);
        *weight = divs3(mul3(scl3(l->albedo, G2), fr), G1);
        return 1;
    }
    if (!refract3(wo, wh, l->ni, l->nts, wi)) return 0;
    if (same_hemi(*wi, wo)) return 0;
    float D = ggx_d(wh, l->ax, l->ay);
    float G1 = ggx_g1(wo, l->ax, l->ay);
    float G2 = ggx_g2(*wi, wo, l->ax, l->ay);
    float dwh = sqrf(l->nts) * fabsf(dot3(*wi, wh)) /
                fmaxr(sqrf(l->ni * dot3(wo, wh) + l->nts * dot3(*wi, wh)), 1E-12f);
    float ini = safe_rcp(l->ni);
    float c = fmaxr(dot3(wo, wh), 0.0f);
    float fr = fresnel_dd(l->nts * ini, c);
    *pdf = D * G1 * fabsf(dot3(wo, wh)) * dwh / fmaxr(fabsf(wo.z), 1E-8f);
    *weight = divs3(scl3(scl3(scl3(l->albedo, 1.0f - fr), sqrf(l->ni / l->nts)), G2), G1);
    return 1;
}

/* ------------------------------------------------------------ Perlin noise */
static const uint8_t perlin_perm[256] = {
#include "../ymrt_amd/csrc/perlin_perm.inc"
};

/* PerlinNoiseEval, src/TextureEval.cuh:55-131 (the 512-entry table holds the
 * 256-entry permutation twice) */
float or_perlin(v3 p)
{
    int x0 = (int)floorf(p.x), y0 = (int)floorf(p.y), z0 = (int)floorf(p.z);
    float tx = p.x - (float)x0, ty = p.y - (float)y0, tz = p.z - (float)z0;
    int x1 = (x0 + 1) & 255, y1 = (y0 + 1) & 255, z1 = (z0 + 1) & 255;
    x0 &= 255; y0 &= 255; z0 &= 255;
#define OR_HASH(x, y, z) perlin_perm[(perlin_perm[(perlin_perm[(x)] + (y)) & 255] + (z)) & 255]
    float g[8];
    int hs[8] = {OR_HASH(x0, y0, z0), OR_HASH(x1, y0, z0), OR_HASH(x0, y1, z0), OR_HASH(x0, y0, z1),
                 OR_HASH(x1, y1, z0), OR_HASH(x0, y1, z1), OR_HASH(x1, y0, z1), OR_HASH(x1, y1, z1)};
#undef OR_HASH
    float xs[8] = {tx, tx - 1, tx, tx, tx - 1, tx, tx - 1, tx - 1};
    float ys[8] = {ty, ty, ty - 1, ty, ty - 1, ty - 1, ty, ty - 1};
    float zs[8] = {tz, tz, tz, tz - 1, tz, tz - 1, tz - 1, tz - 1};
    for (int i = 0; i < 8; ++i) {
        float x = xs[i], y = ys[i], z = zs[i];
        switch (hs[i] & 15) {
        case 0: g[i] = x + y; break;
        case 1: g[i] = -x + y; break;
        case 2: g[i] = x - y; break;
        case 3: g[i] = -x - y; break;
        case 4: g[i] = x + z; break;
        case 5: g[i] = -x + z; break;
        case 6: g[i] = x - z; break;
        case 7: g[i] = -x - z; break;
        case 8: g[i] = y + z; break;
        case 9: g[i] = -y + z; break;
        case 10: g[i] = y - z; break;
        case 11: g[i] = -y - z; break;
        case 12: g[i] = y + x; break;
        case 13: g[i] = -x + y; break;
        case 14: g[i] = -y + z; break;
        default: g[i] = -y - z; break;
        }
    }
    float ix = tx * tx * tx * (tx * (tx * 6.0f - 15.0f) + 10.0f);
    float iy = ty * ty * ty * (ty * (ty * 6.0f - 15.0f) + 10.0f);
    float iz = tz * tz * tz * (tz * (tz * 6.0f - 15.0f) + 10.0f);
    /* g: v000 v100 v010 v001 v110 v011 v101 v111 */
    float v00 = g[0] + ix * (g[1] - g[0]);
    float v01 = g[3] + ix * (g[6] - g[3]);
    float v10 = g[2] + ix * (g[4] - g[2]);
    float v11 = g[5] + ix * (g[7] - g[5]);
    float v0 = v00 + iy * (v10 - v00);
    float v1 = v01 + iy * (v11 - v01);
    return v0 + iz * (v1 - v0);
}

/* PerlinNoiseFBM / PerlinNoiseTurbulence, src/TextureEval.cuh:132-160 */
static float perlin_fbm(v3 p, float lac, float gain, int layers, float amp, float off, int turb)
{
    float r = 0.0f, freq = 1.0f;
    for (int i = 0; i < layers; ++i) {
        float n = or_perlin(sclf3(freq, p));
        r += amp * (turb ? fabsf(n) : n) + off;
        amp *= gain;
        off *= gain;
        freq *= lac;
    }
    return r;
}

static v3 lerp_bw(v3 black, v3 white, float w) { return add3(black, sclf3(w, sub3(white, black))); }

/* EvalTexResult and the EvalTex* nodes, src/TextureEval.cuh:238-413 */
static v3 eval_node(const or_scene *sc, uint32_t ti, v3 black, v3 white, or_tex_coord tc)
{
    const ymrt_texture_desc *t = sc->textures[ti].desc;
    v3 po = tc.p_object;
    switch (t->type) {
    case YMRT_TEXTURE_IMAGE: return or_texture_eval(sc, ti, tc.st.x, tc.st.y);
    case YMRT_TEXTURE_CONSTANT_FLOAT: return s3(t->value[0]);
    case YMRT_TEXTURE_CONSTANT_RGB: return mk3(t->value[0], t->value[1], t->value[2]);
    case YMRT_TEXTURE_CHECKERBOARD: {
        float x = or_sinf(po.x * t->frequency), y = or_sinf(po.y * t->frequency), z = or_sinf(po.z * t->frequency);
        return x * y * z < 0.0f ? white : black;
    }
    case YMRT_TEXTURE_NOISE: {
        float n = or_perlin(scl3(po, t->frequency));
        if (t->normalized) n = n * 0.5f + 0.5f;
        return lerp_bw(black, white, n);
    }
    case YMRT_TEXTURE_FBM:
        return lerp_bw(black, white, perlin_fbm(sclf3(t->frequency, po), t->lacunarity, t->gain, t->layer_count,
                                                 t->amplitude, t->offset, 0));
    case YMRT_TEXTURE_TURBULENCE:
        return lerp_bw(black, white, perlin_fbm(sclf3(t->frequency, po), t->lacunarity, t->gain, t->layer_count,
                                                 t->amplitude, t->offset, 1));
    case YMRT_TEXTURE_MARBLE: {
        float fbm = perlin_fbm(sclf3(t->frequency, po), t->lacunarity, t->gain, t->layer_count, 1.0f, 0.0f, 0);
        float turb = 0.0f;
        turb += (t->turbulence_axes & 1) ? po.x : 0.0f;
        turb += (t->turbulence_axes & 2) ? po.y : 0.0f;
        turb += (t->turbulence_axes & 4) ? po.z : 0.0f;
        float m = or_sinf(turb + t->variation * fbm) * 0.5f + 0.5f;
        return lerp_bw(black, white, m);
    }
    case YMRT_TEXTURE_WOOD: {
        float w = or_perlin(scl3(po, t->frequency)) * t->variation;
        w = w - floorf(w);
        return lerp_bw(black, white, w);
    }
    case YMRT_TEXTURE_POLKA_DOT: {
        float sx = tc.st.x * t->frequency, sy = tc.st.y * t->frequency;
        float fs = floorf(sx), ft = floorf(sy);
        if (!(or_perlin(mk3(fs + 0.5f, ft + 0.5f, 0.0f)) > 0.0f)) return black;
        float hx, hy;
#define OR_FRACT114(f, m) (hx = (f) * (m), (hx - floorf(hx)) * 114.f)
        float max_off = 0.5f - t->radius;
        float ds = sx - fs, dt = sy - ft;
        hy = OR_FRACT114(sx, 1145.14f);
        float a1 = OR_FRACT114(sy, 9807.26f);
        float s_c = 0.5f + max_off * or_perlin(mk3(hy, a1, 0.0f));
        hy = OR_FRACT114(sx, 4444.44f);
        a1 = OR_FRACT114(sy, 2333.33f);
        float t_c = 0.5f + max_off * or_perlin(mk3(hy, a1, 0.0f));
#undef OR_FRACT114
        return sqrf(ds - s_c) + sqrf(dt - t_c) < t->radius * t->radius ? white : black;
    }
    default: { /* WAVE */
        float f0 = perlin_fbm(scl3(po, t->frequency), t->lacunarity, t->gain, t->layer_count, 1.0f, 0.0f, 0);
        float f1 = perlin_fbm(scl3(po, t->frequency1), t->lacunarity1, t->gain1, t->layer_count1, 1.0f, 0.0f, 0);
        return lerp_bw(black, white, fabsf(f0) * f1);
    }
    }
}

/* TextureEval, src/TextureEval.cuh:415-481: an explicit stack of node copies;
 * a node is evaluated once both inputs hold results, which are written into
 * the parent's copy. */
typedef struct {
    uint32_t ti;
    v3 res[2];
    int done[2];
    int parent, ith;
} tex_frame;

v3 or_texture_graph(const or_scene *sc, uint32_t root, or_tex_coord tc)
{
    tex_frame st[8];
    int top = 0;
    memset(st, 0, sizeof(st));
    st[0].ti = root; st[0].parent = -1; st[0].ith = -1;
    for (;;) {
        tex_frame *f = &st[top];
        const ymrt_texture_desc *t = sc->textures[f->ti].desc;
        int nchild = t->type >= YMRT_TEXTURE_CHECKERBOARD ? 2 : 0;
        int next = -1;
        for (int c = 0; c < nchild; ++c) if (!f->done[c]) { next = c; break; }
        if (next < 0) {
            v3 r = eval_node(sc, f->ti, f->res[0], f->res[1], tc);
            if (f->parent < 0) return r;
            st[f->parent].res[f->ith] = r;
            st[f->parent].done[f->ith] = 1;
            --top;
            continue;
        }
        uint32_t ci = t->child[next];
        if (ci == OR_EMPTY) { f->res[next] = s3(0.0f); f->done[next] = 1; continue; }
        if (top + 1 >= 8) return s3(0.0f);
        ++top;
        memset(&st[top], 0, sizeof(tex_frame));
        st[top].ti = ci; st[top].parent = top - 1; st[top].ith = next;
    }
}

static float tex_float(const or_scene *sc, uint32_t tex, float def, or_tex_coord tc)
{
    return tex != OR_EMPTY ? or_texture_graph(sc, tex, tc).x : def;
}
static v3 tex_color(const or_scene *sc, uint32_t tex, const float *def, or_tex_coord tc)
{
    return tex != OR_EMPTY ? or_texture_graph(sc, tex, tc) : mk3(def[0], def[1], def[2]);
}

/* src/BSDF.cuh:453-618 */
static void init_default_bsdf(const or_scene *sc, mbsdf *b, const ymrt_material_desc *m, or_tex_coord uv,
                              v3 ray_d, int hit_back, float ray_ior, float ex_ior)
{
    const float ior_min = 0.0001f;
    float coat_w = tex_float(sc, m->coat_weight_tex, m->coat_weight, uv);
    float ior_in = fmaxr(ray_ior, ior_min);
    float ior_coat = fmaxr(m->coat_ior, ior_min);
    float ior_out = fmaxr(hit_back ? ex_ior : m->ior_n, ior_min);
    float eta_ic = ior_coat * safe_rcp(ior_in);
    float eta_io = ior_out * safe_rcp(ior_in);
    float cos_in = fmaxr(dot3(b->n, neg3(ray_d)), 0.0f);
    float fr_ic = coat_w > 0.0f ? fresnel_dd(eta_ic, cos_in) : 1.0f;
    float fr_io = fresnel_dd(eta_io, cos_in);
    float metal = tex_float(sc, m->metalness_tex, m->metalness, uv);
    float ax = sqrf(clampf(tex_float(sc, m->roughness_x_tex, m->roughness_x, uv), 0.001f, 0.999f));
    float ay = sqrf(clampf(tex_float(sc, m->roughness_y_tex, m->roughness_y, uv), 0.001f, 0.999f));
    float spec_w = tex_float(sc, m->specular_weight_tex, m->specular_weight, uv);
    float trans_w = tex_float(sc, m->transmission_weight_tex, m->transmission_weight, uv);
    v3 dif = tex_color(sc, m->diffuse_albedo_tex, m->diffuse_albedo, uv);
    v3 spe = tex_color(sc, m->specular_albedo_tex, m->specular_albedo, uv);
    float coat_th = tex_float(sc, m->coat_thickness_tex, m->coat_thickness, uv);
    float cax = sqrf(clampf(tex_float(sc, m->coat_roughness_x_tex, m->coat_roughness_x, uv), 0.001f, 0.999f));
    float cay = sqrf(clampf(tex_float(sc, m->coat_roughness_y_tex, m->coat_roughness_y, uv), 0.001f, 0.999f));
    v3 coat_alb = tex_color(sc, m->coat_albedo_tex, m->coat_albedo, uv);
    v3 ca = max3(coat_alb, s3(1E-8f));
    v3 sigma = scl3(neg3(mk3(or_logf(ca.x), or_logf(ca.y), or_logf(ca.z))), coat_th);
    float rc = safe_rcp(cos_in);
    v3 ns = neg3(sigma);
    v3 coat_tr = mk3(or_expf(ns.x * rc), or_expf(ns.y * rc), or_expf(ns.z * rc));
    float wsum = 0.0f;
    b->count = 0;
    v3 coat_color = s3(coat_w * fr_ic);
    float coat_lum = luminance3(coat_color);
    if (coat_lum > 0.0f) {
        lobe *l = &b->l[b->count];
        l->type = LOBE_REFLECT; l->albedo = coat_color; l->ax = cax; l->ay = cay;
        l->nt = s3(ior_coat); l->kt = s3(0.0f); l->ni = ior_in; l->metal = 0;
        b->w[b->count] = coat_lum; wsum += coat_lum; ++b->count;
    }
    /* LinearLerp(vec3(1), coat_tr, coat_w) = (1 - w) * 1 + coat_tr * w */
    v3 lerp = add3(sclf3(1.0f - coat_w, s3(1.0f)), scl3(coat_tr, coat_w));
    v3 tint = sclf3(1.0f - coat_w * fr_ic, lerp);
    v3 metal_c = scl3(tint, metal);
    float metal_lum = luminance3(metal_c);
    if (metal_lum > 0.0f) {
        v3 mn, mk;
        edge_tint(dif, spe, &mn, &mk);
        lobe *l = &b->l[b->count];
        l->type = LOBE_REFLECT; l->albedo = metal_c; l->ax = ax; l->ay = ay;
        l->nt = mn; l->kt = mk; l->ni = ior_in; l->metal = 1;
        b->w[b->count] = metal_lum; wsum += metal_lum; ++b->count;
    }
    v3 spec_c = mul3(scl3(scl3(scl3(tint, 1.0f - metal), fr_io), spec_w), spe);
    float spec_lum = luminance3(spec_c);
    if (spec_lum > 0.0f) {
        lobe *l = &b->l[b->count];
        l->type = LOBE_REFLECT; l->albedo = spec_c; l->ax = ax; l->ay = ay;
        l->nt = s3(ior_out); l->kt = s3(0.0f); l->ni = ior_in; l->metal = 0;
        b->w[b->count] = spec_lum; wsum += spec_lum; ++b->count;
    }
    v3 trans_c = mul3(scl3(scl3(scl3(tint, 1.0f - metal), 1.0f - fr_io * spec_w), trans_w), spe);
    float trans_lum = luminance3(trans_c);
    if (trans_lum > 0.0f) {
        lobe *l = &b->l[b->count];
        l->type = LOBE_TRANSMIT; l->albedo = trans_c; l->ax = ax; l->ay = ay;
        l->ni = ior_in; l->nts = ior_out;
        b->w[b->count] = trans_lum; wsum += trans_lum; ++b->count;
    }
    v3 dif_c = mul3(scl3(scl3(scl3(tint, 1.0f - metal), 1.0f - fr_io * spec_w), 1.0f - trans_w), dif);
    float dif_lum = luminance3(dif_c);
    {
        lobe *l = &b->l[b->count];
        l->type = LOBE_LAMBERT; l->albedo = dif_c;
        b->w[b->count] = dif_lum; wsum += dif_lum; ++b->count;
    }
    if (wsum > 0.0f) {
        float iw = safe_rcp(wsum);
        for (int i = 0; i < b->count; ++i) b->w[i] *= iw;
    } else {
        float iw = 1.0f / (float)b->count;
        for (int i = 0; i < b->count; ++i) b->w[i] = iw;
    }
}

static v3 bsdf_eval(const mbsdf *b, v3 wo, v3 wi, float *pdf)
{
    v3 f = s3(0.0f);
    float p = 0.0f;
    if (!b->is_default) { *pdf = 0.0f; return f; }
    for (int i = 0; i < b->count; ++i) {
        float lp = 0.0f;
        f = add3(f, lobe_eval(&b->l[i], wo, wi, &lp));
        p += b->w[i] * lp;
    }
    *pdf = p;
    return f;
}

static int bsdf_sample(const mbsdf *b, v3 wo, v3 *weight, v3 *wi, float *pdf, float u0, float u1, float u2)
{
    if (!b->is_default) return 0;
    float cdf = 0.0f;
    int sel = -1;
    for (int i = 0; i < b->count; ++i) {
        if (u0 < cdf + b->w[i]) { sel = i; break; }
        cdf += b->w[i];
    }
    if (sel == -1) sel = b->count - 1;
    if (sel == -1) return 0;
    float lp = 0.0f;
    if (!lobe_sample(&b->l[sel], u1, u2, wo, weight, wi, &lp)) return 0;
    float mp = lp * b->w[sel];
    v3 mf = scl3(*weight, lp);
    for (int i = 0; i < b->count; ++i) {
        if (i == sel) continue;
        float p = 0.0f;
        mf = add3(mf, lobe_eval(&b->l[i], wo, *wi, &p));
        mp += b->w[i] * p;
    }
    *pdf = mp;
    *weight = scl3(mf, safe_rcp(mp));
    return 1;
}

static v3 to_local(const mbsdf *b, v3 w) { return mk3(dot3(w, b->t), dot3(w, b->b), dot3(w, b->n)); }
static v3 to_world(const mbsdf *b, v3 w)
{
    return add3(add3(sclf3(w.x, b->t), sclf3(w.y, b->b)), sclf3(w.z, b->n));
}

/* Camera::TextureCoordinatesDifferential, src/Camera.h:56-140: (du/dx, dv/dx,
 * du/dy, dv/dy) from rays through neighbouring pixels offset by (sx, sy). */
static void camera_uv_differentials(const or_scene *sc, v3 hp, v3 sn, v3 ng, v3 dpdu, v3 dpdv, uint32_t width,
                                    uint32_t height, float sx, float sy, float *duv)
{
    v3 look = sub3(mk3(sc->camera.position[0], sc->camera.position[1], sc->camera.position[2]), hp);
    float dist = length3(look);
    look = sclf3(dist != 0.0f ? 1.0f / dist : 0.0f, look);
    v3 right = normalize3(sub3(sc->cam_u, sclf3(dot3(sc->cam_u, look), look)));
    v3 up = cross3(look, right);
    float iph = 2.0f * sc->tan_half_fov;
    float yoff = iph * sy / (float)height;
    float ipw = iph * sc->camera.aspect_ratio;
    float xoff = ipw * sx / (float)width;
    v3 dxd = normalize3(add3(mk3(0.0f, 0.0f, -1.0f), sclf3(xoff, right)));
    v3 dyd = normalize3(add3(mk3(0.0f, 0.0f, -1.0f), sclf3(yoff, up)));
    v3 lsn = mk3(dot3(sn, right), dot3(sn, up), dot3(sn, look));
    v3 lp = mk3(0.0f, 0.0f, -dist);
    float ddx = dot3(lsn, dxd), ddy = dot3(lsn, dyd);
    if (fabsf(ddx) < 1E-4f) ddx = ddx > 0.0f ? 1E-4f : -1E-4f;
    if (fabsf(ddy) < 1E-4f) ddy = ddy > 0.0f ? 1E-4f : -1E-4f;
    float tx = dot3(lsn, lp) / ddx, ty = dot3(lsn, lp) / ddy;
    v3 ex = sclf3(sx != 0.0f ? 1.0f / sx : 0.0f, sub3(sclf3(tx, dxd), lp));
    v3 ey = sclf3(sy != 0.0f ? 1.0f / sy : 0.0f, sub3(sclf3(ty, dyd), lp));
    v3 dpdx = add3(add3(sclf3(ex.x, right), sclf3(ex.y, up)), sclf3(ex.z, look));
    v3 dpdy = add3(add3(sclf3(ey.x, right), sclf3(ey.y, up)), sclf3(ey.z, look));
    float a[3] = {fabsf(ng.x), fabsf(ng.y), fabsf(ng.z)};
    int d0 = 1, d1 = 2;
    if (a[2] > a[1] && a[2] > a[0]) { d0 = 0; d1 = 1; }
    else if (a[1] > a[0]) { d0 = 0; d1 = 2; }
    float U[3] = {dpdu.x, dpdu.y, dpdu.z}, V[3] = {dpdv.x, dpdv.y, dpdv.z};
    float X[3] = {dpdx.x, dpdx.y, dpdx.z}, Y[3] = {dpdy.x, dpdy.y, dpdy.z};
    float m00 = U[d0], m01 = V[d0], m10 = U[d1], m11 = V[d1];
    float det = m00 * m11 - m10 * m01;
    if (fabsf(det) < 1E-7f) { duv[0] = duv[1] = duv[2] = duv[3] = 0.0f; return; }
    float idet = 1.0f / det;
    duv[0] = (m11 * X[d0] - m01 * X[d1]) * idet;
    duv[1] = (-m10 * X[d0] + m00 * X[d1]) * idet;
    duv[2] = (m11 * Y[d0] - m01 * Y[d1]) * idet;
    duv[3] = (-m10 * Y[d0] + m00 * Y[d1]) * idet;
}

static void frame_from(mbsdf *b, int hit_back, v3 sn, v3 su)
{
    if (!hit_back) {
        b->n = sn; b->t = normalize3(su); b->b = cross3(b->n, b->t);
    } else {
        b->n = neg3(sn); b->t = neg3(normalize3(su)); b->b = neg3(cross3(b->n, b->t));
    }
}

/* MaterialBSDF::InitShadingSpace, src/BSDF.cuh:694-807: shading frame from the
 * interpolated normal and dpdu, perturbed by a normal map (tangent-space
 * normal in [0,1]^3) or a bump map (forward differences of the height over
 * the texture-coordinate footprint), then reflected to face the ray. */
static void init_shading_space(const or_scene *sc, mbsdf *b, const ymrt_material_desc *m, int hit_back, v3 ray_d,
                               v3 ng, v3 sn_in, v3 dpdu, v3 dpdv, const float *w2o, or_tex_coord tc,
                               const float *duv)
{
    uint32_t ntex = m ? m->normal_mapping_tex : OR_EMPTY;
    uint32_t btex = m ? m->bump_mapping_tex : OR_EMPTY;
    v3 sn = sn_in;
    if (dot3(sn_in, ng) < 0.0f) sn = neg3(sn_in);
    v3 su = dpdu;
    if (dot3(cross3(dpdu, dpdv), ng) < 0.0f) su = neg3(dpdu);
    v3 sv = dpdv;
    su = sub3(su, scl3(sn, dot3(su, sn)));
    sv = sub3(sv, scl3(sn, dot3(sv, sn)));
    if (ntex != OR_EMPTY) {
        frame_from(b, hit_back, sn, su);
        v3 tn = normalize3(sub3(scl3(or_texture_graph(sc, ntex, tc), 2.0f), s3(1.0f)));
        b->n = add3(add3(sclf3(tn.x, b->t), sclf3(tn.y, b->b)), sclf3(tn.z, b->n));
        b->t = normalize3(sub3(b->t, scl3(b->n, dot3(b->t, b->n))));
        b->b = hit_back ? neg3(cross3(b->n, b->t)) : cross3(b->n, b->t);
    } else if (btex != OR_EMPTY) {
        v3 osu = transform_vector(w2o, su), osv = transform_vector(w2o, sv);
        float du = fmaxr((fabsf(duv[0]) + fabsf(duv[2])) * 0.5f, 0.0005f);
        if (!(du > 0.0f)) du = 0.0005f;
        or_tex_coord t1 = tc;
        t1.st = mk2(tc.st.x + du, tc.st.y);
        t1.p_world = add3(tc.p_world, scl3(su, du));
        t1.p_object = add3(tc.p_object, scl3(osu, du));
        float hu = or_texture_graph(sc, btex, t1).x;
        float dv = fmaxr((fabsf(duv[1]) + fabsf(duv[3])) * 0.5f, 0.0005f);
        if (!(dv > 0.0f)) dv = 0.0005f;
        or_tex_coord t2 = tc;
        t2.st = mk2(tc.st.x, tc.st.y + dv);
        t2.p_world = add3(tc.p_world, scl3(sv, dv));
        t2.p_object = add3(tc.p_object, scl3(osv, dv));
        float hv = or_texture_graph(sc, btex, t2).x;
        float h0 = or_texture_graph(sc, btex, tc).x;
        su = add3(su, sclf3((hu - h0) * (du != 0.0f ? 1.0f / du : 0.0f), sn));
        sv = add3(sv, sclf3((hv - h0) * (dv != 0.0f ? 1.0f / dv : 0.0f), sn));
        sn = normalize3(cross3(su, sv));
        frame_from(b, hit_back, sn, su);
    } else {
        frame_from(b, hit_back, sn, su);
    }
    if (dot3(neg3(ray_d), b->n) < 0.0f) {
        b->n = reflect3(b->n, dot3(b->n, ng) > 0.0f ? ng : neg3(ng));
        b->t = normalize3(sub3(b->t, scl3(b->n, dot3(b->t, b->n))));
        b->b = hit_back ? neg3(cross3(b->n, b->t)) : cross3(b->n, b->t);
    }
}

int or_bsdf_eval(const ymrt_material_desc *m, v3 wo, v3 wi, float ray_ior, float ex_ior, int hit_back,
                 float *f_rgb, float *pdf)
{
    mbsdf b;
    memset(&b, 0, sizeof(b));
    b.t = mk3(1, 0, 0); b.b = mk3(0, 1, 0); b.n = mk3(0, 0, 1);
    b.is_default = m->type == YMRT_MATERIAL_DEFAULT;
    or_scene dummy;
    memset(&dummy, 0, sizeof(dummy));
    or_tex_coord tc0;
    memset(&tc0, 0, sizeof(tc0));
    if (b.is_default) init_default_bsdf(&dummy, &b, m, tc0, neg3(wo), hit_back, ray_ior, ex_ior);
    v3 f = bsdf_eval(&b, wo, wi, pdf);
    f_rgb[0] = f.x; f_rgb[1] = f.y; f_rgb[2] = f.z;
    return b.count;
}

int or_bsdf_sample(const ymrt_material_desc *m, v3 wo, float ray_ior, float ex_ior, int hit_back,
                   float u0, float u1, float u2, float *w_rgb, float *wi_out, float *pdf)
{
    mbsdf b;
    memset(&b, 0, sizeof(b));
    b.t = mk3(1, 0, 0); b.b = mk3(0, 1, 0); b.n = mk3(0, 0, 1);
    b.is_default = m->type == YMRT_MATERIAL_DEFAULT;
    or_scene dummy;
    memset(&dummy, 0, sizeof(dummy));
    or_tex_coord tc0;
    memset(&tc0, 0, sizeof(tc0));
    if (b.is_default) init_default_bsdf(&dummy, &b, m, tc0, neg3(wo), hit_back, ray_ior, ex_ior);
    v3 w = s3(0.0f), wi = s3(0.0f);
    *pdf = 0.0f;
    int ok = bsdf_sample(&b, wo, &w, &wi, pdf, u0, u1, u2);
    w_rgb[0] = w.x; w_rgb[1] = w.y; w_rgb[2] = w.z;
    wi_out[0] = wi.x; wi_out[1] = wi.y; wi_out[2] = wi.z;
    return ok;
}

/* ------------------------------------------------------------------ phases */
float or_hg_eval(float g, v3 wo, v3 wi, float *pdf)
{
    float c = dot3(wo, wi);
    float t = 1.f + g * g + 2.f * g * c;
    *pdf = 0.25f * OR_INV_PI * (1.f - g * g) * safe_rcp(t * sqrtf(fmaxr(t, 0.0f)));
    return *pdf;
}

int or_hg_sample(float g, float u0, float u1, v3 wo, v3 *wi, float *pdf)
{
    float c;
    if (fabsf(g) < 1e-4f) {
        c = 1.f - 2.f * u0;
    } else {
        float t = (1.f - g * g) * safe_rcp(1.f - g + 2.f * g * u0);
        c = (1.f + g * g - t * t) / (2.f * g);
    }
    float s = sqrtf(fmaxr(0.0f, 1.0f - c * c));
    float phi = 2.f * OR_PI * u1;
    v3 w = wo, u, v;
    orthogonal_basis(&w, &u, &v);
    *wi = add3(add3(scl3(scl3(u, s), or_cosf(phi)), scl3(scl3(v, s), or_sinf(phi))), scl3(w, c));
    or_hg_eval(g, wo, *wi, pdf);
    return 1;
}

static int mixed_phase_sample(const float *vw, const float *gs, int n, float u0, float u1, v3 wo,
                              float *weight, v3 *wi, float *pdf)
{
    if (n == 0) return 0;
    float cdf = 0.0f;
    int sel = -1;
    for (int i = 0; i < n; ++i) {
        if (u0 < cdf + vw[i]) { sel = i; u0 = (u0 - cdf) * safe_rcp(vw[i]); break; }
        cdf += vw[i];
    }
    if (sel == -1) sel = n - 1;
    float pp = 0.0f;
    or_hg_sample(gs[sel], u0, u1, wo, wi, &pp);
    float w1 = 1.0f;
    float mp = pp * vw[sel];
    float mg = w1 * pp * vw[sel];
    for (int i = 0; i < n; ++i) {
        if (i == sel) continue;
        float p;
        mg += vw[i] * or_hg_eval(gs[i], wo, *wi, &p);
        mp += vw[i] * p;
    }
    *pdf = mp;
    *weight = mg * safe_rcp(mp);
    return 1;
}

static float mixed_phase_eval(const float *vw, const float *gs, int n, v3 wo, v3 wi, float *pdf)
{
    if (n == 0) { *pdf = 0.0f; return 0.0f; }
    float mg = 0.0f, mp = 0.0f;
    for (int i = 0; i < n; ++i) {
        float p = 0.0f;
        float g = or_hg_eval(gs[i], wo, wi, &p);
        mg += vw[i] * g;
        mp += vw[i] * p;
    }
    *pdf = mp;
    return mg;
}

/* ------------------------------------------------------------------ volumes */
typedef struct {
    or_sampler *s;        /* path sampler (free-flight sampling)            */
    uint32_t hash_key[4]; /* shadow transmittance stream for grid media     */
    uint32_t hash_ctr;
} rng_ctx;

/* Path-replay adjoint state of one path (or_render_adjoint).  Gradient
 * terms come in two kinds: a "downstream" event (param, w, mark) adds
 * adj . (L_path - mark) * w to the parameter once the path's radiance is
 * known -- the factor it stands for multiplies every deposit after the mark
 * -- and a direct term adds adj . deposit * w for one shadow ray's deposit
 * (the BSDF and the transmittance along that ray).  Sampling pdfs and MIS
 * weights are held fixed (detached): the estimator differentiates the
 * integrand only, which is unbiased for the derivative of the expected
 * image. */
#define OR_ADJ_RAY_EVENTS 1024
typedef struct { uint32_t param; v3 w; v3 mark; } adj_down;
typedef struct { uint32_t param; float w; } adj_ev;
typedef struct {
    const or_adj_targets *tg;
    adj_down *down;
    uint32_t nd, cap;
    adj_ev ray_ev[OR_ADJ_RAY_EVENTS];
    uint32_t nray;
    int recording;      /* shadow-ray transmittance events are wanted */
    double *grad;
    v3 adj;             /* the pixel's dLoss/dImage / spp */
    v3 beta;            /* throughput at the current vertex */
    v3 L;               /* radiance so far (mark of free-flight events) */
    int rough_param;    /* the surface vertex's roughness parameter, or -1 */
    float rough_r;      /* and its roughness */
} or_adj_path;

static void adj_add_down(or_adj_path *a, uint32_t param, v3 w, v3 mark)
{
    if (a->nd == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 256;
        a->down = (adj_down *)realloc(a->down, sizeof(adj_down) * a->cap);
    }
    adj_down e = {param, w, mark};
    a->down[a->nd++] = e;
}

/* the voxel of a grid holding object point p, or -1 outside the unit cube
 * (grid_density's cell) */
static int grid_voxel(const or_scene *sc, int gi, v3 p)
{
    if (gi < 0 || (uint32_t)gi >= sc->grid_count) return -1;
    const or_grid *g = &sc->grids[gi];
    if (!(p.x >= 0.0f && p.x < 1.0f && p.y >= 0.0f && p.y < 1.0f && p.z >= 0.0f && p.z < 1.0f)) return -1;
    uint32_t ix = (uint32_t)(p.x * (float)g->nx), iy = (uint32_t)(p.y * (float)g->ny), iz = (uint32_t)(p.z * (float)g->nz);
    if (ix >= g->nx) ix = g->nx - 1;
    if (iy >= g->ny) iy = g->ny - 1;
    if (iz >= g->nz) iz = g->nz - 1;
    return (int)(((size_t)iz * g->ny + iy) * g->nx + ix);
}

static float shadow_rand(rng_ctx *r)
{
    uint32_t v[4] = {r->hash_key[0], r->hash_key[1] ^ (r->hash_ctr << 8), r->hash_key[2], r->hash_key[3]};
    r->hash_ctr++;
    or_pcg4d(v);
    return (float)(v[0] >> 8) * (1.0f / 16777216.0f);
}

static float grid_density(const or_scene *sc, int gi, v3 p)
{
    const or_grid *g = &sc->grids[gi];
    if (!(p.x >= 0.0f && p.x < 1.0f && p.y >= 0.0f && p.y < 1.0f && p.z >= 0.0f && p.z < 1.0f)) return 0.0f;
    uint32_t ix = (uint32_t)(p.x * (float)g->nx), iy = (uint32_t)(p.y * (float)g->ny), iz = (uint32_t)(p.z * (float)g->nz);
    if (ix >= g->nx) ix = g->nx - 1;
    if (iy >= g->ny) iy = g->ny - 1;
    if (iz >= g->nz) iz = g->nz - 1;
    return g->density[((size_t)iz * g->ny + iy) * g->nx + ix];
}

/* heterogeneous: a voxel grid, or a density texture as the reference's
 * density_texture_idx (src/VolumeEval.cuh:231-238) */
static int vol_is_het(const or_scene *sc, int v)
{
    int gi = sc->volumes[v].density_grid;
    int ti = sc->volumes[v].density_texture;
    return (gi >= 0 && (uint32_t)gi < sc->grid_count) || (ti >= 0 && (uint32_t)ti < sc->texture_count);
}

/* Texture::GetUpperBound (src/Texture.h:1181-1251) of the graph's root */
static float texture_upper_bound(const ymrt_texture_desc *t)
{
    switch (t->type) {
    case YMRT_TEXTURE_IMAGE: return 0.0f;
    case YMRT_TEXTURE_CONSTANT_FLOAT: return t->value[0];
    case YMRT_TEXTURE_CONSTANT_RGB: return fmaxr(t->value[0], fmaxr(t->value[1], t->value[2]));
    case YMRT_TEXTURE_FBM:
    case YMRT_TEXTURE_TURBULENCE: {
        float result = 0.0f, amplitude = t->amplitude, offset = t->offset;
        for (int i = 0; i < t->layer_count; ++i) {
            result += amplitude + offset;
            amplitude *= t->gain;
            offset *= t->gain;
        }
        return result;
    }
    case YMRT_TEXTURE_WAVE: return safe_rcp(1.0f - t->gain) * safe_rcp(1.0f - t->gain1);
    default: return 1.0f;  /* checkerboard, noise, marble, wood, polka dot */
    }
}

/* majorant of a heterogeneous volume's density */
static float vol_upper_bound(const or_scene *sc, const ymrt_volume_desc *v)
{
    if (v->density_texture >= 0) return texture_upper_bound(sc->textures[v->density_texture].desc);
    return sc->grids[v->density_grid].max_density;
}

/* density at a tracking step, max(., 0): the texture's x at (st = 0, world
 * point, object point) as src/VolumeEval.cuh:165-176, or the grid voxel */
static float vol_density(const or_scene *sc, const ymrt_volume_desc *v, v3 pw, v3 po)
{
    if (v->density_texture >= 0) {
        or_tex_coord tc;
        tc.st = mk2(0.0f, 0.0f);
        tc.p_world = pw;
        tc.p_object = po;
        return fmaxr(or_texture_graph(sc, (uint32_t)v->density_texture, tc).x, 0.0f);
    }
    return fmaxr(grid_density(sc, v->density_grid, po), 0.0f);
}

static v3 vsig(const ymrt_volume_desc *v) { return mk3(v->sigma_t[0], v->sigma_t[1], v->sigma_t[2]); }
static v3 valb(const ymrt_volume_desc *v) { return mk3(v->albedo[0], v->albedo[1], v->albedo[2]); }

/* src/VolumeEval.cuh:18-78 */
static int sample_homogeneous(const or_scene *sc, int n, const int *vi, float tmax, v3 *trw, float *dist,
                              float *vw, or_sampler *s)
{
    *trw = s3(1.0f);
    *dist = OR_FLOAT_MAX;
    v3 st = s3(0.0f), ss = s3(0.0f);
    for (int i = 0; i < n; ++i) {
        const ymrt_volume_desc *v = &sc->volumes[vi[i]];
        st = add3(st, vsig(v));
        ss = add3(ss, mul3(vsig(v), valb(v)));
    }
    float csum = st.x + st.y + st.z;
    v3 cw = csum > 0.0f ? divs3(st, csum) : s3(1.0f / 3.0f);
    float u0 = or_sampler_next(s);
    int ch;
    if (u0 < cw.x) { ch = 0; u0 = u0 * safe_rcp(cw.x); }
    else if (u0 < cw.x + cw.y) { ch = 1; u0 = (u0 - cw.x) * safe_rcp(cw.y); }
    else { ch = 2; u0 = (u0 - cw.x - cw.y) * safe_rcp(cw.z); }
    float sc_ = comp3(st, ch);
    float fpl = sc_ > 0.0f ? (-or_logf(1.0f - u0) / sc_) : OR_FLOAT_MAX;
    int scatter = fpl < tmax;
    float m = fminr(tmax, fpl);
    v3 tr = mk3(or_expf(-st.x * m), or_expf(-st.y * m), or_expf(-st.z * m));
    float pdf = scatter ? (cw.x * st.x * tr.x + cw.y * st.y * tr.y + cw.z * st.z * tr.z)
                        : (cw.x * tr.x + cw.y * tr.y + cw.z * tr.z);
    if (!(fabsf(pdf) > 0.0f)) return 0;
    float w = 1.0f / (float)n;
    for (int i = 0; i < n; ++i) vw[i] = w;
    *trw = scatter ? divs3(mul3(tr, ss), pdf) : divs3(tr, pdf);
    *dist = fpl;
    return scatter;
}

/* src/VolumeEval.cuh:80-214, densities from the voxel grid or a texture graph */
static int sample_heterogeneous(const or_scene *sc, int n, const int *vi, const or_ray *ray, float tmax,
                                v3 *trw, float *dist, float *vw, or_sampler *s, or_adj_path *adj)
{
    *trw = s3(1.0f);
    *dist = OR_FLOAT_MAX;
    v3 ro[OR_MAX_BOUNDARY_RECORD], rd[OR_MAX_BOUNDARY_RECORD];
    v3 sm = s3(0.0f);
    for (int i = 0; i < n; ++i) {
        const ymrt_volume_desc *v = &sc->volumes[vi[i]];
        float ub = 1.0f;
        if (vol_is_het(sc, vi[i])) {
            ro[i] = m4_point(v->world_to_object, ray->o);
            rd[i] = m4_dir(v->world_to_object, ray->d);
            ub = vol_upper_bound(sc, v);
        }
        sm = add3(sm, scl3(vsig(v), ub));
    }
    float csum = sm.x + sm.y + sm.z;
    v3 cw = csum > 0.0f ? divs3(sm, csum) : s3(1.0f / 3.0f);
    float u0 = or_sampler_next(s);
    int ch = u0 < cw.x ? 0 : (u0 < cw.x + cw.y ? 1 : 2);
    if (!(comp3(sm, ch) > 0.0f)) return 0;
    v3 nf = mk3(safe_rcp(sm.x + 1.0f), safe_rcp(sm.y + 1.0f), safe_rcp(sm.z + 1.0f));
    v3 smr = mk3(safe_rcp(sm.x), safe_rcp(sm.y), safe_rcp(sm.z));
    v3 tr = s3(1.0f), pdf = s3(1.0f);
    float acc = 0.0f;
    int it = 0;
    for (;;) {
        if (it >= OR_MAX_FREE_PATH_LENGTH) break;
        float fpl = -or_logf(1.0f - or_sampler_next(s)) / comp3(sm, ch);
        if (acc + fpl > tmax) {
            float dd = tmax - acc;
            v3 ptr = mk3(or_expf(-sm.x * dd), or_expf(-sm.y * dd), or_expf(-sm.z * dd));
            tr = mul3(tr, mul3(ptr, nf));
            pdf = mul3(pdf, mul3(ptr, nf));
            acc += fpl;
            break;
        }
        acc += fpl;
        v3 st = s3(0.0f), ss = s3(0.0f);
        int tvox = -1;
        float tsig = 0.0f;
        for (int i = 0; i < n; ++i) {
            const ymrt_volume_desc *v = &sc->volumes[vi[i]];
            float dens = 1.0f;
            v3 po = add3(ro[i], scl3(rd[i], acc));
            if (vol_is_het(sc, vi[i])) dens = vol_density(sc, v, add3(ray->o, scl3(ray->d, acc)), po);
            st = add3(st, scl3(vsig(v), dens));
            ss = add3(ss, scl3(mul3(vsig(v), valb(v)), dens));
            if (adj && (uint32_t)vi[i] == adj->tg->vol) {
                tvox = grid_voxel(sc, v->density_grid, po);
                tsig = v->sigma_t[0];
            }
        }
        v3 rp;
        rp.x = fminr(sm.x > 0.0f ? st.x * smr.x : 0.0f, 1.0f);
        rp.y = fminr(sm.y > 0.0f ? st.y * smr.y : 0.0f, 1.0f);
        rp.z = fminr(sm.z > 0.0f ? st.z * smr.z : 0.0f, 1.0f);
        v3 fp = sub3(s3(1.0f), rp);
        v3 ptr = mk3(or_expf(-sm.x * fpl), or_expf(-sm.y * fpl), or_expf(-sm.z * fpl));
        if (or_sampler_next(s) < comp3(rp, ch)) {
            tr = mul3(tr, mul3(mul3(ptr, ss), nf));
            pdf = mul3(pdf, mul3(mul3(mul3(ptr, sm), rp), nf));
            /* a real collision: probability sigma(x) / sigma_m (gray media) */
            if (tvox >= 0 && st.x > 0.0f)
                adj_add_down(adj, (uint32_t)(adj->tg->off_vox + (size_t)tvox), s3(tsig / st.x), adj->L);
            break;
        }
        /* a null collision: probability 1 - sigma(x) / sigma_m */
        if (tvox >= 0 && sm.x - st.x > 0.0f)
            adj_add_down(adj, (uint32_t)(adj->tg->off_vox + (size_t)tvox), s3(-tsig / (sm.x - st.x)), adj->L);
        tr = mul3(tr, mul3(mul3(mul3(ptr, sm), fp), nf));
        pdf = mul3(pdf, mul3(mul3(mul3(ptr, sm), fp), nf));
        ++it;
    }
    float w = 1.0f / (float)n;
    for (int i = 0; i < n; ++i) vw[i] = w;
    *trw = scl3(tr, safe_rcp(dot3(cw, pdf)));
    *dist = acc;
    return (acc > tmax || it >= OR_MAX_FREE_PATH_LENGTH) ? 0 : 1;
}

static int sample_volume(const or_scene *sc, const or_ray *ray, int n, const int *vi, float tmax,
                         v3 *trw, float *dist, float *vw, or_sampler *s, or_adj_path *adj)
{
    if (sc->volume_count == 0 || n == 0) return 0;
    int het = 0;
    for (int i = 0; i < n; ++i) if (vol_is_het(sc, vi[i])) { het = 1; break; }
    if (het) return sample_heterogeneous(sc, n, vi, ray, tmax, trw, dist, vw, s, adj);
    return sample_homogeneous(sc, n, vi, tmax, trw, dist, vw, s);
}

/* src/VolumeEval.cuh:253-385 */
static v3 eval_tr(const or_scene *sc, const or_ray *ray, const or_transfer *t, float tmax, rng_ctx *rng,
                  or_adj_path *adj)
{
    if (sc->volume_count == 0) return s3(1.0f);
    int vi[OR_MAX_BOUNDARY_RECORD];
    int n = tr_volumes(t, vi, sc->volume_count);
    if (n == 0) return s3(1.0f);
    int het = 0;
    for (int i = 0; i < n; ++i) if (vol_is_het(sc, vi[i])) { het = 1; break; }
    if (!het) {
        v3 st = s3(0.0f);
        for (int i = 0; i < n; ++i) st = add3(st, vsig(&sc->volumes[vi[i]]));
        return mk3(or_expf(-st.x * tmax), or_expf(-st.y * tmax), or_expf(-st.z * tmax));
    }
    v3 ro[OR_MAX_BOUNDARY_RECORD], rd[OR_MAX_BOUNDARY_RECORD];
    v3 sm = s3(0.0f);
    for (int i = 0; i < n; ++i) {
        const ymrt_volume_desc *v = &sc->volumes[vi[i]];
        float ub = 1.0f;
        if (vol_is_het(sc, vi[i])) {
            ro[i] = m4_point(v->world_to_object, ray->o);
            rd[i] = m4_dir(v->world_to_object, ray->d);
            ub = vol_upper_bound(sc, v);
        }
        sm = add3(sm, scl3(vsig(v), ub));
    }
    float csum = sm.x + sm.y + sm.z;
    v3 cw = csum > 0.0f ? divs3(sm, csum) : s3(1.0f / 3.0f);
    float u0 = shadow_rand(rng);
    int ch = u0 < cw.x ? 0 : (u0 < cw.x + cw.y ? 1 : 2);
    if (!(comp3(sm, ch) > 0.0f)) return s3(1.0f);
    v3 nf = mk3(safe_rcp(sm.x + 1.0f), safe_rcp(sm.y + 1.0f), safe_rcp(sm.z + 1.0f));
    v3 smr = mk3(safe_rcp(sm.x), safe_rcp(sm.y), safe_rcp(sm.z));
    v3 tr = s3(1.0f), pdf = s3(1.0f);
    float acc = 0.0f;
    int it = 0;
    for (;;) {
        if (it >= OR_MAX_FREE_PATH_LENGTH) break;
        float fpl = -or_logf(1.0f - shadow_rand(rng)) / comp3(sm, ch);
        if (acc + fpl > tmax) {
            float dd = tmax - acc;
            v3 ptr = mk3(or_expf(-sm.x * dd), or_expf(-sm.y * dd), or_expf(-sm.z * dd));
            tr = mul3(tr, mul3(ptr, nf));
            pdf = mul3(pdf, mul3(ptr, nf));
            break;
        }
        acc += fpl;
        v3 st = s3(0.0f);
        int tvox = -1;
        float tsig = 0.0f;
        for (int i = 0; i < n; ++i) {
            const ymrt_volume_desc *v = &sc->volumes[vi[i]];
            float dens = 1.0f;
            v3 po = add3(ro[i], scl3(rd[i], acc));
            if (vol_is_het(sc, vi[i])) dens = vol_density(sc, v, add3(ray->o, scl3(ray->d, acc)), po);
            st = add3(st, scl3(vsig(v), dens));
            if (adj && adj->recording && (uint32_t)vi[i] == adj->tg->vol) {
                tvox = grid_voxel(sc, v->density_grid, po);
                tsig = v->sigma_t[0];
            }
        }
        /* ratio tracking's factor 1 - sigma(x) / sigma_m (gray media) */
        if (tvox >= 0 && sm.x - st.x > 0.0f && adj->nray < OR_ADJ_RAY_EVENTS) {
            adj_ev e = {(uint32_t)(adj->tg->off_vox + (size_t)tvox), -tsig / (sm.x - st.x)};
            adj->ray_ev[adj->nray++] = e;
        }
        v3 rp;
        rp.x = fminr(sm.x > 0.0f ? st.x * smr.x : 0.0f, 1.0f);
        rp.y = fminr(sm.y > 0.0f ? st.y * smr.y : 0.0f, 1.0f);
        rp.z = fminr(sm.z > 0.0f ? st.z * smr.z : 0.0f, 1.0f);
        v3 fp = sub3(s3(1.0f), rp);
        v3 ptr = mk3(or_expf(-sm.x * fpl), or_expf(-sm.y * fpl), or_expf(-sm.z * fpl));
        tr = mul3(tr, mul3(mul3(mul3(ptr, sm), fp), nf));
        pdf = mul3(pdf, mul3(mul3(ptr, sm), nf));
        if (it + 1 > 32) {
            v3 pt = scl3(tr, safe_rcp(dot3(cw, pdf)));
            float rr = fminr(fmaxr(pt.x, fmaxr(pt.y, pt.z)), 1.0f);
            if (shadow_rand(rng) < rr) tr = divs3(tr, fmaxr(rr, OR_MIN_COLOR_EPSILON));
            else break;
        }
        ++it;
    }
    return scl3(tr, safe_rcp(dot3(cw, pdf)));
}

/* src/VolumeEval.cuh:420-468 */
static v3 trace_tr(const or_scene *sc, const or_ray *ray, or_transfer *t, rng_ctx *rng, uint64_t *counts,
                   or_adj_path *adj)
{
    if (sc->volume_count == 0) return s3(1.0f);
    v3 tr = s3(1.0f);
    or_ray r = *ray;
    float maxd = ray->t;
    float traced = OR_TMIN;
    while (traced < maxd) {
        or_hit h;
        memset(&h, 0, sizeof(h));
        h.t = OR_TMAX; h.inst = OR_EMPTY; h.tri = OR_EMPTY;
        or_nearby nb[OR_MAX_NEARBY];
        int nc = 0;
        if (counts) counts[3]++;
        int hb = or_trace_volume(sc, &r, &h, &nc, nb);
        if (!hb) {
            tr = mul3(tr, eval_tr(sc, &r, t, fmaxr(maxd - traced, 0.0f), rng, adj));
            break;
        }
        tr = mul3(tr, eval_tr(sc, &r, t, h.t, rng, adj));
        traced += h.t;
        v3 p, perr, ng;
        or_fetch_geometry_normal(sc, &h, &p, &perr, &ng);
        boundary_bunch(sc, t, nb, nc);
        r = make_ray(or_offset_ray_origin(p, perr, r.d, ng), r.d, h.inst, h.tri);
    }
    return tr;
}

/* ------------------------------------------------------------------ lights */
static int sample_light_power(const or_scene *sc, float u, float *pdf)
{
    int left = 0, right = (int)sc->light_count;
    while (left < right) {
        int mid = (left + right) / 2;
        if (sc->light_cdf[mid] <= u) left = mid + 1; else right = mid;
    }
    int idx = left;
    if (idx > (int)sc->light_count - 1) idx = (int)sc->light_count - 1;
    *pdf = idx == 0 ? sc->light_cdf[idx] : (sc->light_cdf[idx] - sc->light_cdf[idx - 1]);
    return idx;
}

typedef struct {
    v3 p0, p1, p2, n0, n1, n2;
    const ymrt_instance_desc *in;
    const ymrt_material_desc *mtl;
} light_tri;

static void load_light(const or_scene *sc, const or_shape_light *l, light_tri *lt)
{
    lt->in = &sc->instances[l->instance];
    lt->mtl = &sc->materials[lt->in->material];
    const or_mesh *m = &sc->meshes[lt->in->geometry];
    uint32_t t = l->triangle;
    lt->p0 = ld3(m->pos, m->pidx[3 * t]);
    lt->p1 = ld3(m->pos, m->pidx[3 * t + 1]);
    lt->p2 = ld3(m->pos, m->pidx[3 * t + 2]);
    if (m->nidx && m->nrm) {
        lt->n0 = ld3(m->nrm, m->nidx[3 * t]);
        lt->n1 = ld3(m->nrm, m->nidx[3 * t + 1]);
        lt->n2 = ld3(m->nrm, m->nidx[3 * t + 2]);
    } else {
        v3 g = cross3(sub3(lt->p1, lt->p0), sub3(lt->p2, lt->p0));
        lt->n0 = g; lt->n1 = g; lt->n2 = g;
    }
}

static v3 light_le(const light_tri *lt)
{
    return scl3(mk3(lt->mtl->light_color[0], lt->mtl->light_color[1], lt->mtl->light_color[2]), lt->mtl->intensity);
}

static v3 interp3(v3 a, v3 b, v3 c, v3 w) { return add3(add3(scl3(a, w.x), scl3(b, w.y)), scl3(c, w.z)); }

/* RayIntersectSphere, src/Light.h:42-65 (4.0 is a double literal) */
static int ray_sphere_light(v3 o, v3 d, float radius, float *t0, float *t1)
{
    float a = d.x * d.x + d.y * d.y + d.z * d.z;
    float b = 2.0f * (d.x * o.x + d.y * o.y + d.z * o.z);
    float c = o.x * o.x + o.y * o.y + o.z * o.z - radius * radius;
    float disc = (float)((double)(b * b) - 4.0 * (double)a * (double)c);
    if (disc < 0.0f) return 0;
    float sq = sqrtf(disc);
    float ia = 1.0f / a;
    *t0 = (-b - sq) * 0.5f * ia;
    *t1 = (-b + sq) * 0.5f * ia;
    if (*t1 < 1E-10f) return 0;
    return 1;
}

/* SampleCone, src/SampleUtilities.h:33-39 */
static v3 sample_cone(float u0, float u1, float ctm)
{
    float ct = (1.0f - u0) + u0 * ctm;
    float st = sqrtf(fmaxr(0.0f, 1.0f - ct * ct));
    float phi = u1 * 2.0f * OR_PI;
    return mk3(or_cosf(phi) * st, or_sinf(phi) * st, ct);
}

static v3 inst_le(const or_scene *sc, const ymrt_instance_desc *in)
{
    const ymrt_material_desc *m = &sc->materials[in->material];
    return scl3(mk3(m->light_color[0], m->light_color[1], m->light_color[2]), m->intensity);
}

static v3 m4_translate(const float *m) { return mk3(m[12], m[13], m[14]); }

/* src/Light.h:282-317 (sphere) */
static int sphere_light_sample_li(const or_scene *sc, const or_shape_light *l, v3 pos, float u0, float u1, v3 *wi,
                                  v3 *lpos, v3 *lerr, v3 *lng, v3 *color, float *pdf)
{
    const ymrt_instance_desc *in = &sc->instances[l->instance];
    v3 opos = transform_position(in->world_to_object, pos, NULL);
    float lpo = length3(opos);
    float radius = sc->spheres[in->geometry].radius;
    if (lpo < radius) { *pdf = 0.0f; return 0; }
    float ctm = sqrtf(fmaxr(sqrf(lpo) - sqrf(radius), 0.0f)) / lpo;
    float cone = 2.0f * OR_PI * (1.0f - ctm);
    v3 cs = sample_cone(u0, u1, ctm);
    v3 w = neg3(opos), u, v;
    orthogonal_basis(&w, &u, &v);
    cs = add3(add3(sclf3(cs.x, u), sclf3(cs.y, v)), sclf3(cs.z, w));
    float t0 = 0.0f, t1 = 0.0f;
    ray_sphere_light(opos, cs, radius, &t0, &t1);
    v3 ip = add3(opos, scl3(cs, t0));
    float lio = length3(ip);
    ip = scl3(ip, radius * safe_rcp(lio));
    v3 ie = sclf3(error_gamma(5), abs3(ip));
    *lpos = transform_position(in->object_to_world, ip, &ie);
    *lerr = ie;
    *lng = normalize3(sub3(*lpos, m4_translate(in->object_to_world)));
    *wi = normalize3(transform_vector(in->object_to_world, cs));
    *pdf = 1.0f / cone;
    *color = scl3(inst_le(sc, in), cone);
    return 1;
}

/* src/Light.h:163-199 (sphere) */
static v3 sphere_light_eval_li(const or_scene *sc, const or_shape_light *l, v3 pos, v3 wi, v3 *lpos, v3 *lerr,
                               v3 *lng, float *pdf)
{
    const ymrt_instance_desc *in = &sc->instances[l->instance];
    v3 owi = transform_vector(in->world_to_object, wi);
    v3 opos = transform_position(in->world_to_object, pos, NULL);
    float lpo = length3(opos);
    float radius = sc->spheres[in->geometry].radius;
    if (lpo < radius) { *pdf = 0.0f; return s3(0.0f); }
    float t0, t1;
    if (!ray_sphere_light(opos, owi, radius, &t0, &t1)) { *pdf = 0.0f; return s3(0.0f); }
    float ctm = sqrtf(fmaxr(sqrf(lpo) - sqrf(radius), 0.0f)) / lpo;
    float cone = 2.0f * OR_PI * (1.0f - ctm);
    float ct = dot3(normalize3(owi), divs3(neg3(opos), lpo));
    if (ct < ctm) { *pdf = 0.0f; return s3(0.0f); }
    v3 ip = add3(opos, scl3(owi, t0));
    float lio = length3(ip);
    ip = scl3(ip, radius * safe_rcp(lio));
    v3 ie = sclf3(error_gamma(5), abs3(ip));
    *lpos = transform_position(in->object_to_world, ip, &ie);
    *lerr = ie;
    *lng = normalize3(sub3(*lpos, m4_translate(in->object_to_world)));
    *pdf = 1.0f / cone;
    return inst_le(sc, in);
}

/* src/Light.h:207-281 (mesh) */
static int light_sample_li(const or_scene *sc, const or_shape_light *l, v3 pos, float u0, float u1, v3 *wi,
                           v3 *lpos, v3 *lerr, v3 *lng, v3 *color, float *pdf)
{
    if (l->triangle == OR_EMPTY) return sphere_light_sample_li(sc, l, pos, u0, u1, wi, lpos, lerr, lng, color, pdf);
    light_tri lt;
    load_light(sc, l, &lt);
    v3 gn = cross3(sub3(lt.p1, lt.p0), sub3(lt.p2, lt.p0));
    float su = sqrtf(fmaxr(0.0f, u0));
    float bu = 1.0f - su, bv = u1 * su;
    v3 b = mk3(bu, bv, 1.0f - bu - bv);
    float xs = fabsf(lt.p0.x * b.x) + fabsf(lt.p1.x * b.y) + fabsf(lt.p2.x * b.z);
    float ys = fabsf(lt.p0.y * b.x) + fabsf(lt.p1.y * b.y) + fabsf(lt.p2.y * b.z);
    float zs = fabsf(lt.p0.z * b.x) + fabsf(lt.p1.z * b.y) + fabsf(lt.p2.z * b.z);
    v3 sp = interp3(lt.p0, lt.p1, lt.p2, b);
    v3 se = sclf3(error_gamma(7), mk3(xs, ys, zs));
    v3 sn = interp3(lt.n0, lt.n1, lt.n2, b);
    v3 wn = normalize3(m4_transpose_dir(lt.in->world_to_object, sn));
    *lpos = transform_position(lt.in->object_to_world, sp, &se);
    *lerr = se;
    *lng = normalize3(transpose_transform_vector(lt.in->world_to_object, gn));
    float dist = length3(sub3(*lpos, pos));
    *wi = scl3(sub3(*lpos, pos), safe_rcp(dist));
    float ct = -dot3(wn, *wi);
    if (ct <= 0.0f) { *pdf = 0.0f; return 0; }
    *pdf = sqrf(dist) * safe_rcp(l->area * ct);
    *color = divs3(light_le(&lt), *pdf);
    return 1;
}

/* src/Light.h:14-40 with glm 0.9.8's mat3 inverse (func_matrix.inl:272-294) */
static int ray_tri_light(v3 o, v3 d, v3 p0, v3 p1, v3 p2, v3 *bary)
{
    v3 a = sub3(p0, p2), b = sub3(p1, p2);
    float m[3][3];
    m[0][0] = a.x; m[0][1] = a.y; m[0][2] = a.z;
    m[1][0] = b.x; m[1][1] = b.y; m[1][2] = b.z;
    m[2][0] = -d.x; m[2][1] = -d.y; m[2][2] = -d.z;
    float ood = 1.0f / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
                        - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])
                        + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
    float iv[3][3];
    iv[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * ood;
    iv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * ood;
    iv[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * ood;
    iv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * ood;
    iv[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * ood;
    iv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * ood;
    iv[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * ood;
    iv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * ood;
    iv[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * ood;
    v3 r = sub3(o, p2);
    float u = iv[0][0] * r.x + iv[1][0] * r.y + iv[2][0] * r.z;
    float v = iv[0][1] * r.x + iv[1][1] * r.y + iv[2][1] * r.z;
    float t = iv[0][2] * r.x + iv[1][2] * r.y + iv[2][2] * r.z;
    float w = 1.0f - u - v;
    if (t < OR_TMIN) return 0;
    *bary = mk3(u, v, w);
    return (u > 0.0f && u < 1.0f) && (v > 0.0f && v < 1.0f) && (w > 0.0f && w < 1.0f);
}

/* src/Light.h:87-162 (mesh) */
static v3 light_eval_li(const or_scene *sc, const or_shape_light *l, v3 pos, v3 wi, v3 *lpos, v3 *lerr,
                        v3 *lng, float *pdf)
{
    if (l->triangle == OR_EMPTY) return sphere_light_eval_li(sc, l, pos, wi, lpos, lerr, lng, pdf);
    light_tri lt;
    load_light(sc, l, &lt);
    v3 gn = cross3(sub3(lt.p1, lt.p0), sub3(lt.p2, lt.p0));
    v3 owi = transform_vector(lt.in->world_to_object, wi);
    v3 opos = transform_position(lt.in->world_to_object, pos, NULL);
    v3 b = s3(0.0f);
    if (!ray_tri_light(opos, owi, lt.p0, lt.p1, lt.p2, &b)) { *pdf = 0.0f; return s3(0.0f); }
    float xs = fabsf(lt.p0.x * b.x) + fabsf(lt.p1.x * b.y) + fabsf(lt.p2.x * b.z);
    float ys = fabsf(lt.p0.y * b.x) + fabsf(lt.p1.y * b.y) + fabsf(lt.p2.y * b.z);
    float zs = fabsf(lt.p0.z * b.x) + fabsf(lt.p1.z * b.y) + fabsf(lt.p2.z * b.z);
    v3 sp = interp3(lt.p0, lt.p1, lt.p2, b);
    v3 se = sclf3(error_gamma(7), mk3(xs, ys, zs));
    v3 sn = interp3(lt.n0, lt.n1, lt.n2, b);
    v3 wn = normalize3(m4_transpose_dir(lt.in->world_to_object, sn));
    float ct = -dot3(wn, wi);
    if (ct <= 0.0f) { *pdf = 0.0f; return s3(0.0f); }
    *lpos = transform_position(lt.in->object_to_world, sp, &se);
    *lerr = se;
    *lng = normalize3(transpose_transform_vector(lt.in->world_to_object, gn));
    float dist = length3(sub3(*lpos, pos));
    *pdf = sqrf(dist) * safe_rcp(l->area * ct);
    return light_le(&lt);
}

static float power_heuristic(float main_pdf, float vice_pdf)
{
    return safe_rcp(1.0f + sqrf(vice_pdf * safe_rcp(main_pdf)));
}

static v3 clampv(v3 v) { return min3(v, s3(OR_MAX_COLOR_CLAMP)); }

typedef struct {
    const or_scene *sc;
    const ymrt_render_settings *rs;
    or_sampler *s;
    rng_ctx *rng;
    uint64_t *counts;
    or_adj_path *adj;
} path_ctx;

/* one shadow ray's deposit, beta * term * scale: its roughness term (g, at a
 * surface vertex on a target material) and the transmittance events
 * recorded along it */
static void adj_nee(path_ctx *c, v3 term, float scale, int rough_param, float g)
{
    or_adj_path *a = c->adj;
    if (!a) return;
    v3 dep = sclf3(scale, mul3(a->beta, term));
    double d = (double)a->adj.x * dep.x + (double)a->adj.y * dep.y + (double)a->adj.z * dep.z;
    if (rough_param >= 0) a->grad[rough_param] += d * (double)g;
    for (uint32_t i = 0; i < a->nray; ++i) a->grad[a->ray_ev[i].param] += d * (double)a->ray_ev[i].w;
    a->nray = 0;
}

/* DistantLight::SampleLi (src/Light.h:361-385) */
static int distant_sample_li(const or_scene *sc, uint32_t i, float u0, float u1, v3 *wi, v3 *color, float *pdf)
{
    const ymrt_distant_light_desc *dl = &sc->distant[i];
    float ctm = sc->distant_cos_max[i];
    float cone = 2.0f * OR_PI * (1.0f - ctm);
    v3 cs;
    if (cone < 1E-10f) {
        cs = mk3(0.0f, 1.0f, 0.0f);
    } else {
        cs = sample_cone(u0, u1, ctm);
        cs = mk3(cs.x, cs.z, cs.y);
    }
    *wi = neg3(normalize3(transform_vector(dl->object_to_world, cs)));
    *pdf = 1.0f / fmaxr(cone, 1E-10f);
    *color = scl3(mk3(dl->light_color[0], dl->light_color[1], dl->light_color[2]), dl->intensity);
    return 1;
}

/* DistantLight::EvalLi (src/Light.h:344-358): radiance times the cone pdf */
static v3 distant_eval_li(const or_scene *sc, uint32_t i, v3 wi, float *pdf)
{
    const ymrt_distant_light_desc *dl = &sc->distant[i];
    v3 lw = transform_vector(dl->world_to_object, wi);
    float ct = -lw.y;
    float ctm = sc->distant_cos_max[i];
    if (ct < ctm) { *pdf = 0.0f; return s3(0.0f); }
    float cone = 2.0f * OR_PI * (1.0f - ctm);
    *pdf = 1.0f / cone;
    return scl3(scl3(mk3(dl->light_color[0], dl->light_color[1], dl->light_color[2]), dl->intensity), *pdf);
}

static uint32_t pick_distant(const or_scene *sc, float u)
{
    int k = (int)(u * (float)sc->distant_count);
    if (k > (int)sc->distant_count - 1) k = (int)sc->distant_count - 1;
    if (k < 0) k = 0;
    return (uint32_t)k;
}

/* d ln f / d r of a metal-only material's GGX lobe at (wo, wi) (local),
 * f = albedo F D G2 / (4 wo.z) with alpha = clamp(r, 0.001, 0.999)^2:
 * d ln D / d alpha + d ln G2 / d alpha, times 2 r (0 where the clamp holds).
 * Isotropic GGX: D = 1 / (pi a^2 (s / a^2 + hz^2)^2) with s = hx^2 + hy^2,
 * Lambda(w) = (sqrt(1 + a^2 t^2) - 1) / 2 with t^2 = (wx^2 + wy^2) / wz^2. */
static float rough_dlog(float r, v3 wo, v3 wi)
{
    if (!(r > 0.001f && r < 0.999f) || !(wo.z > 0.0f && wi.z > 0.0f)) return 0.0f;
    double a = (double)r * r;
    double hx = (double)wo.x + wi.x, hy = (double)wo.y + wi.y, hz = (double)wo.z + wi.z;
    double hl = sqrt(hx * hx + hy * hy + hz * hz);
    if (hl < 1e-12) return 0.0f;
    hx /= hl; hy /= hl; hz /= hl;
    double sh = hx * hx + hy * hy, q = sh / (a * a) + hz * hz;
    double dlnd = -2.0 / a + 4.0 * sh / (a * a * a * q);
    double lam[2] = {0.0, 0.0}, dlam[2] = {0.0, 0.0};
    v3 ws[2] = {wo, wi};
    for (int k = 0; k < 2; ++k) {
        double wz = ws[k].z;
        if (fabs(wz) < 1e-8) continue;
        double t2 = ((double)ws[k].x * ws[k].x + (double)ws[k].y * ws[k].y) / (wz * wz);
        double root = sqrt(1.0 + a * a * t2);
        lam[k] = 0.5 * (root - 1.0);
        dlam[k] = 0.5 * a * t2 / root;
    }
    double dlng = -(dlam[0] + dlam[1]) / (1.0 + lam[0] + lam[1]);
    return (float)((dlnd + dlng) * 2.0 * r);
}

static int adj_rough_param(const path_ctx *c) { return c->adj ? c->adj->rough_param : -1; }
static float adj_rough_g(const path_ctx *c, v3 wo, v3 wi)
{
    return c->adj && c->adj->rough_param >= 0 ? rough_dlog(c->adj->rough_r, wo, wi) : 0.0f;
}

/* EvalDistantLight, src/Shading.cu:83-198 */
static v3 eval_distant_light(path_ctx *c, mbsdf *b, v3 ray_d, const or_hit *h, const or_transfer *t,
                             int nc, const or_nearby *nb, v3 pos, v3 perr, v3 ng)
{
    const or_scene *sc = c->sc;
    if (sc->distant_count == 0) return s3(0.0f);
    v3 direct = s3(0.0f);
    v3 wo = to_local(b, neg3(ray_d));
    float sel_pdf = 1.0f / (float)sc->distant_count;
    uint32_t li = pick_distant(sc, or_sampler_next(c->s));
    {
        v3 ldir = s3(0.0f), Li = s3(0.0f);
        float lpdf = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        int valid = distant_sample_li(sc, li, u0, u1, &ldir, &Li, &lpdf);
        v3 lwi = to_local(b, ldir);
        float wpdf = 0.0f;
        v3 f = bsdf_eval(b, wo, lwi, &wpdf);
        if (maxcomp3(f) > OR_MIN_COLOR_EPSILON && valid) {
            v3 so = or_offset_ray_origin(pos, perr, ldir, ng);
            or_ray sr = make_ray(so, ldir, OR_EMPTY, OR_EMPTY);
            int transmit = dot3(neg3(ray_d), ng) * dot3(ldir, ng) < 0.0f;
            int invalid = (!same_hemi(wo, lwi) && !transmit) || (same_hemi(wo, lwi) && transmit);
            if (!invalid) {
                if (c->counts) c->counts[2]++;
                if (!or_trace_shadow(sc, &sr)) {
                    v3 tr = s3(1.0f);
                    if (c->rs->enable_volume_scattering) {
                        or_transfer st = *t;
                        if (transmit) boundary_bunch(sc, &st, nb, nc);
                        sr = make_ray(so, ldir, h->inst, h->tri);
                        tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                    }
                    float mis = power_heuristic(lpdf, wpdf);
                    v3 term = mul3(mul3(sclf3(mis, tr), clampv(f)), clampv(Li));
                    direct = add3(direct, term);
                    adj_nee(c, term, 1.0f / sel_pdf, adj_rough_param(c), adj_rough_g(c, wo, lwi));
                }
            }
        }
    }
    {
        v3 bw = s3(0.0f), bwi = s3(0.0f);
        float bpdf = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        float u2 = or_sampler_next(c->s);
        if (bsdf_sample(b, wo, &bw, &bwi, &bpdf, u0, u1, u2)) {
            v3 ldir = to_world(b, bwi);
            float lpdf = 0.0f;
            v3 Li = distant_eval_li(sc, li, ldir, &lpdf);
            if (maxcomp3(Li) > OR_MIN_COLOR_EPSILON) {
                v3 so = or_offset_ray_origin(pos, perr, ldir, ng);
                or_ray sr = make_ray(so, ldir, OR_EMPTY, OR_EMPTY);
                int transmit = dot3(neg3(ray_d), ng) * dot3(ldir, ng) < 0.0f;
                int invalid = (!same_hemi(wo, bwi) && !transmit) || (same_hemi(wo, bwi) && transmit);
                if (!invalid) {
                    if (c->counts) c->counts[2]++;
                    if (!or_trace_shadow(sc, &sr)) {
                        v3 tr = s3(1.0f);
                        if (c->rs->enable_volume_scattering) {
                            or_transfer st = *t;
                            if (transmit) boundary_bunch(sc, &st, nb, nc);
                            sr = make_ray(so, ldir, h->inst, h->tri);
                            tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                        }
                        float mis = power_heuristic(bpdf, lpdf);
                        v3 term = mul3(mul3(sclf3(mis, tr), clampv(bw)), clampv(Li));
                        direct = add3(direct, term);
                        adj_nee(c, term, 1.0f / sel_pdf, adj_rough_param(c), adj_rough_g(c, wo, bwi));
                    }
                }
            }
        }
    }
    return divs3(direct, sel_pdf);
}

/* src/Shading.cu:294-417 */
static v3 eval_shape_light(path_ctx *c, mbsdf *b, v3 ray_d, const or_hit *h, const or_transfer *t,
                           int nc, const or_nearby *nb, v3 pos, v3 perr, v3 ng)
{
    const or_scene *sc = c->sc;
    if (sc->light_count == 0) return s3(0.0f);
    v3 direct = s3(0.0f);
    v3 wo = to_local(b, neg3(ray_d));
    float sel_pdf = 0.0f;
    int li = sample_light_power(sc, or_sampler_next(c->s), &sel_pdf);
    const or_shape_light *l = &sc->lights[li];
    {
        v3 ldir = s3(0.0f), lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f), Li = s3(0.0f);
        float lpdf = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        int valid = light_sample_li(sc, l, pos, u0, u1, &ldir, &lpos, &lerr, &lng, &Li, &lpdf);
        v3 lwi = to_local(b, ldir);
        float wpdf = 0.0f;
        v3 f = bsdf_eval(b, wo, lwi, &wpdf);
        if (maxcomp3(f) > OR_MIN_COLOR_EPSILON && valid) {
            v3 so = or_offset_ray_origin(pos, perr, ldir, ng);
            or_ray sr = make_ray(so, ldir, OR_EMPTY, OR_EMPTY);
            float maxd = length3(sub3(or_offset_ray_origin(lpos, lerr, neg3(ldir), lng), pos)) * OR_SHADOW_RAY_CLAMP;
            sr.t = maxd;
            int transmit = dot3(neg3(ray_d), ng) * dot3(ldir, ng) < 0.0f;
            int invalid = (!same_hemi(wo, lwi) && !transmit) || (same_hemi(wo, lwi) && transmit);
            if (!invalid) {
                if (c->counts) c->counts[2]++;
                if (!or_trace_shadow(sc, &sr)) {
                    v3 tr = s3(1.0f);
                    if (c->rs->enable_volume_scattering) {
                        or_transfer st = *t;
                        if (transmit) boundary_bunch(sc, &st, nb, nc);
                        sr = make_ray(so, ldir, h->inst, h->tri);
                        sr.t = maxd;
                        tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                    }
                    float mis = power_heuristic(lpdf, wpdf);
                    v3 term = mul3(mul3(sclf3(mis, tr), clampv(f)), clampv(Li));
                    direct = add3(direct, term);
                    adj_nee(c, term, safe_rcp(sel_pdf), adj_rough_param(c), adj_rough_g(c, wo, lwi));
                }
            }
        }
    }
    {
        v3 bw = s3(0.0f), bwi = s3(0.0f);
        float bpdf = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        float u2 = or_sampler_next(c->s);
        if (bsdf_sample(b, wo, &bw, &bwi, &bpdf, u0, u1, u2)) {
            v3 ldir = to_world(b, bwi);
            v3 lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f);
            float lpdf = 0.0f;
            v3 Li = light_eval_li(sc, l, pos, ldir, &lpos, &lerr, &lng, &lpdf);
            if (maxcomp3(Li) > OR_MIN_COLOR_EPSILON) {
                v3 so = or_offset_ray_origin(pos, perr, ldir, ng);
                or_ray sr = make_ray(so, ldir, OR_EMPTY, OR_EMPTY);
                float maxd = length3(sub3(or_offset_ray_origin(lpos, lerr, neg3(ldir), lng), pos)) * OR_SHADOW_RAY_CLAMP;
                sr.t = maxd;
                int transmit = dot3(neg3(ray_d), ng) * dot3(ldir, ng) < 0.0f;
                int invalid = (!same_hemi(wo, bwi) && !transmit) || (same_hemi(wo, bwi) && transmit);
                if (!invalid) {
                    if (c->counts) c->counts[2]++;
                    if (!or_trace_shadow(sc, &sr)) {
                        v3 tr = s3(1.0f);
                        if (c->rs->enable_volume_scattering) {
                            or_transfer st = *t;
                            if (transmit) boundary_bunch(sc, &st, nb, nc);
                            sr = make_ray(so, ldir, h->inst, h->tri);
                            sr.t = maxd;
                            tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                        }
                        float mis = power_heuristic(bpdf, lpdf);
                        v3 term = mul3(mul3(sclf3(mis, tr), clampv(bw)), clampv(Li));
                        direct = add3(direct, term);
                        adj_nee(c, term, safe_rcp(sel_pdf), adj_rough_param(c), adj_rough_g(c, wo, bwi));
                    }
                }
            }
        }
    }
    return scl3(direct, safe_rcp(sel_pdf));
}

/* EvalVolumeDistantLight, src/Shading.cu:202-290 */
static v3 eval_volume_distant(path_ctx *c, int n, const float *gs, const float *vw, const or_ray *ray,
                              const or_transfer *t, v3 vpos)
{
    const or_scene *sc = c->sc;
    if (sc->distant_count == 0 || n == 0) return s3(0.0f);
    v3 direct = s3(0.0f);
    float sel_pdf = 1.0f / (float)sc->distant_count;
    uint32_t li = pick_distant(sc, or_sampler_next(c->s));
    {
        v3 ldir = s3(0.0f), Li = s3(0.0f);
        float lpdf = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        int valid = distant_sample_li(sc, li, u0, u1, &ldir, &Li, &lpdf);
        float wpdf = 0.0f;
        float ph = mixed_phase_eval(vw, gs, n, neg3(ray->d), ldir, &wpdf);
        if (ph > OR_MIN_COLOR_EPSILON && valid) {
            or_ray sr = make_ray(vpos, ldir, OR_EMPTY, OR_EMPTY);
            if (c->counts) c->counts[2]++;
            if (!or_trace_shadow(sc, &sr)) {
                or_transfer st = *t;
                sr = make_ray(vpos, ldir, OR_EMPTY, OR_EMPTY);
                v3 tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                float mis = power_heuristic(lpdf, wpdf);
                v3 term = mul3(scl3(sclf3(mis, tr), fminr(ph, OR_MAX_COLOR_CLAMP)), clampv(Li));
                direct = add3(direct, term);
                adj_nee(c, term, sel_pdf > 0.0f ? 1.0f / sel_pdf : 0.0f, -1, 0.0f);
            }
        }
    }
    {
        v3 pwi = s3(0.0f);
        float ppdf = 0.0f, pw = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        if (mixed_phase_sample(vw, gs, n, u0, u1, neg3(ray->d), &pw, &pwi, &ppdf)) {
            float lpdf = 0.0f;
            v3 Li = distant_eval_li(sc, li, pwi, &lpdf);
            if (maxcomp3(Li) > OR_MIN_COLOR_EPSILON) {
                or_ray sr = make_ray(vpos, pwi, OR_EMPTY, OR_EMPTY);
                if (c->counts) c->counts[2]++;
                if (!or_trace_shadow(sc, &sr)) {
                    or_transfer st = *t;
                    sr = make_ray(vpos, pwi, OR_EMPTY, OR_EMPTY);
                    v3 tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                    float mis = power_heuristic(ppdf, lpdf);
                    v3 term = mul3(scl3(sclf3(mis, tr), fminr(pw, OR_MAX_COLOR_CLAMP)), clampv(Li));
                    direct = add3(direct, term);
                    adj_nee(c, term, sel_pdf > 0.0f ? 1.0f / sel_pdf : 0.0f, -1, 0.0f);
                }
            }
        }
    }
    return divs3(direct, sel_pdf);
}

/* src/Shading.cu:419-513 */
static v3 eval_volume_light(path_ctx *c, int n, const float *gs, const float *vw, const or_ray *ray,
                            const or_transfer *t, v3 vpos)
{
    const or_scene *sc = c->sc;
    if (sc->light_count == 0 || n == 0) return s3(0.0f);
    v3 direct = s3(0.0f);
    float sel_pdf = 0.0f;
    int li = sample_light_power(sc, or_sampler_next(c->s), &sel_pdf);
    const or_shape_light *l = &sc->lights[li];
    {
        v3 ldir = s3(0.0f), lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f), Li = s3(0.0f);
        float lpdf = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        int valid = light_sample_li(sc, l, vpos, u0, u1, &ldir, &lpos, &lerr, &lng, &Li, &lpdf);
        float wpdf = 0.0f;
        float ph = mixed_phase_eval(vw, gs, n, neg3(ray->d), ldir, &wpdf);
        if (ph > OR_MIN_COLOR_EPSILON && valid) {
            or_ray sr = make_ray(vpos, ldir, OR_EMPTY, OR_EMPTY);
            float maxd = length3(sub3(or_offset_ray_origin(lpos, lerr, neg3(ldir), lng), vpos)) * OR_SHADOW_RAY_CLAMP;
            sr.t = maxd;
            if (c->counts) c->counts[2]++;
            if (!or_trace_shadow(sc, &sr)) {
                or_transfer st = *t;
                sr = make_ray(vpos, ldir, OR_EMPTY, OR_EMPTY);
                sr.t = maxd;
                v3 tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                float mis = power_heuristic(lpdf, wpdf);
                v3 term = mul3(scl3(sclf3(mis, tr), fminr(ph, OR_MAX_COLOR_CLAMP)), clampv(Li));
                direct = add3(direct, term);
                adj_nee(c, term, sel_pdf > 0.0f ? 1.0f / sel_pdf : 0.0f, -1, 0.0f);
            }
        }
    }
    {
        v3 pwi = s3(0.0f);
        float ppdf = 0.0f, pw = 0.0f;
        float u0 = or_sampler_next(c->s);
        float u1 = or_sampler_next(c->s);
        if (mixed_phase_sample(vw, gs, n, u0, u1, neg3(ray->d), &pw, &pwi, &ppdf)) {
            v3 ldir = pwi;
            v3 lpos = s3(0.0f), lerr = s3(0.0f), lng = s3(0.0f);
            float lpdf = 0.0f;
            v3 Li = light_eval_li(sc, l, vpos, ldir, &lpos, &lerr, &lng, &lpdf);
            if (maxcomp3(Li) > OR_MIN_COLOR_EPSILON) {
                or_ray sr = make_ray(vpos, ldir, OR_EMPTY, OR_EMPTY);
                float maxd = length3(sub3(or_offset_ray_origin(lpos, lerr, neg3(ldir), lng), vpos)) * OR_SHADOW_RAY_CLAMP;
                sr.t = maxd;
                if (c->counts) c->counts[2]++;
                if (!or_trace_shadow(sc, &sr)) {
                    or_transfer st = *t;
                    sr = make_ray(vpos, ldir, OR_EMPTY, OR_EMPTY);
                    sr.t = maxd;
                    v3 tr = trace_tr(sc, &sr, &st, c->rng, c->counts, c->adj);
                    float mis = power_heuristic(ppdf, lpdf);
                    v3 term = mul3(scl3(sclf3(mis, tr), fminr(pw, OR_MAX_COLOR_CLAMP)), clampv(Li));
                    direct = add3(direct, term);
                    adj_nee(c, term, sel_pdf > 0.0f ? 1.0f / sel_pdf : 0.0f, -1, 0.0f);
                }
            }
        }
    }
    return sel_pdf > 0.0f ? divs3(direct, sel_pdf) : s3(0.0f);
}

static v3 background(v3 d)
{
    return mix3(s3(1.0f), mk3(0.05f, 0.2f, 0.85f), d.y * 0.5f + 0.5f);
}

static v3 light_clamp(const ymrt_render_settings *rs, v3 c, int depth)
{
    if (!(rs->enable_indirect_clamp && depth > 1)) return c;
    if (!(rs->indirect_limit > 0.0f)) return s3(0.0f);
    float lum = luminance3(c);
    if (lum <= 1E-10f) return c;
    float lim = rs->indirect_limit;
    /* glibc's float powf, as the library's float powf here */
    float pw = fminr(0.05f + 0.25f * powf(lum / lim, 1.5f), 1.0f);
    float sf = powf(lim / (lim + lum), pw);
    return scl3(c, sf);
}

/* FetchMeshShadingData + world transform (src/ShadingUtilities.cuh:713-875) */
static void fetch_shading(const or_scene *sc, const or_hit *h, v3 *p, v3 *perr, v3 *sn, v3 *ng, v2 *uv,
                          v3 *dpdu, v3 *dpdv, v3 *p_obj)
{
    const ymrt_instance_desc *in = &sc->instances[h->inst];
    if (in->geometry_type == YMRT_GEOMETRY_SPHERE) {
        /* FetchSphereShadingData, src/ShadingUtilities.cuh:636-711 */
        float radius = fmaxr(sc->spheres[in->geometry].radius, OR_FLOAT_EPSILON);
        v3 po = h->bary;
        v3 no = divs3(po, radius);
        *p = po;
        *perr = sclf3(error_gamma(5), abs3(po));
        *sn = no;
        *ng = no;
        float theta = or_acosf(no.y);
        float phi = or_atan2f(no.z, no.x);
        if (phi < 0.0f) phi += OR_TWO_PI;
        uv->x = 0.5f * phi * OR_INV_PI;
        uv->y = theta * OR_INV_PI;
        float sin_theta = 0.0f;
        if (1.0f - no.y * no.y > 0.0f) sin_theta = sqrtf(1.0f - no.y * no.y);
        float cos_phi = 0.0f, sin_phi = 0.0f;
        if (sin_theta > 0.0f) { cos_phi = no.x / sin_theta; sin_phi = no.z / sin_theta; }
        else { cos_phi = 1.0f; sin_phi = 0.0f; }
        *dpdu = scl3(mk3(-po.z, 0.0f, po.x), OR_TWO_PI);
        *dpdv = scl3(mk3(po.y * cos_phi, -radius * sin_theta, sin_phi * po.y), OR_PI);
        *p_obj = *p;
        *p = transform_position(in->object_to_world, *p, perr);
        *sn = normalize3(transpose_transform_vector(in->world_to_object, *sn));
        *ng = normalize3(transpose_transform_vector(in->world_to_object, *ng));
        *dpdu = transform_vector(in->object_to_world, *dpdu);
        *dpdv = transform_vector(in->object_to_world, *dpdv);
        return;
    }
    const or_mesh *m = &sc->meshes[in->geometry];
    uint32_t t = h->tri;
    v3 p0 = ld3(m->pos, m->pidx[3 * t]), p1 = ld3(m->pos, m->pidx[3 * t + 1]), p2 = ld3(m->pos, m->pidx[3 * t + 2]);
    v3 b = h->bary;
    float xs = fabsf(p0.x * b.x) + fabsf(p1.x * b.y) + fabsf(p2.x * b.z);
    float ys = fabsf(p0.y * b.x) + fabsf(p1.y * b.y) + fabsf(p2.y * b.z);
    float zs = fabsf(p0.z * b.x) + fabsf(p1.z * b.y) + fabsf(p2.z * b.z);
    *p = interp3(p0, p1, p2, b);
    *perr = sclf3(error_gamma(7), mk3(xs, ys, zs));
    *ng = cross3(sub3(p1, p0), sub3(p2, p0));
    if (m->nidx && m->nrm) {
        *sn = interp3(ld3(m->nrm, m->nidx[3 * t]), ld3(m->nrm, m->nidx[3 * t + 1]), ld3(m->nrm, m->nidx[3 * t + 2]), b);
    } else {
        *sn = *ng;
    }
    int has_uv = m->tidx != NULL && m->uv != NULL;
    v2 uv0 = has_uv ? mk2(m->uv[2 * m->tidx[3 * t]], m->uv[2 * m->tidx[3 * t] + 1]) : mk2(0.0f, 0.0f);
    v2 uv1 = has_uv ? mk2(m->uv[2 * m->tidx[3 * t + 1]], m->uv[2 * m->tidx[3 * t + 1] + 1]) : mk2(1.0f, 0.0f);
    v2 uv2 = has_uv ? mk2(m->uv[2 * m->tidx[3 * t + 2]], m->uv[2 * m->tidx[3 * t + 2] + 1]) : mk2(1.0f, 1.0f);
    uv->x = uv0.x * b.x + uv1.x * b.y + uv2.x * b.z;
    uv->y = uv0.y * b.x + uv1.y * b.y + uv2.y * b.z;
    float du01 = uv1.x - uv0.x, dv01 = uv1.y - uv0.y;
    float du02 = uv2.x - uv0.x, dv02 = uv2.y - uv0.y;
    float det = du01 * dv02 - dv01 * du02;
    if (fabsf(det) < 1E-8f) { du01 = 1.0f; dv01 = 0.0f; du02 = 0.0f; dv02 = 1.0f; det = 1.0f; }
    float idet = 1.0f / det;
    v3 dp01 = sub3(p1, p0), dp02 = sub3(p2, p0);
    *dpdu = scl3(sub3(sclf3(dv02, dp01), sclf3(dv01, dp02)), idet);
    *dpdv = scl3(add3(sclf3(-du02, dp01), sclf3(du01, dp02)), idet);
    *p_obj = *p;
    *p = transform_position(in->object_to_world, *p, perr);
    *sn = normalize3(transpose_transform_vector(in->world_to_object, *sn));
    *ng = normalize3(transpose_transform_vector(in->world_to_object, *ng));
    *dpdu = transform_vector(in->object_to_world, *dpdu);
    *dpdv = transform_vector(in->object_to_world, *dpdv);
}

/* ------------------------------------------------------------- path tracer */
/* ray generation: src/Shading.cu:1634-1725 */
static or_ray primary_ray(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t px, uint32_t py,
                          uint32_t sample, or_sampler *sp)
{
    or_sampler s;
    memset(&s, 0, sizeof(s));
    v2 off;
    uint32_t pixel_index = py * rs->width + px;
    if (rs->sampler_type == OR_SAMPLER_HALTON) {
        or_enumerator e;
        or_enum_init(&e, (int)rs->width, (int)rs->height);
        uint64_t idx = or_enum_index(&e, (int)px, (int)py, (frame - 1) * rs->spp + (int)sample);
        s.type = OR_SAMPLER_HALTON; s.index = idx; s.dim = 2; s.table = sc->halton;
        float ox = or_radical_inverse2(idx), oy = or_radical_inverse3(idx);
        ox = fmaf(ox, e.x_scale, -(float)px);
        oy = fmaf(oy, e.y_scale, -(float)py);
        off.x = clampf(ox, 0.000001f, 0.999999f);
        off.y = clampf(oy, 0.000001f, 0.999999f);
    } else {
        s.type = OR_SAMPLER_PCG; s.pixel = pixel_index; s.sample = sample; s.dim = 0; s.frame = (uint32_t)(frame - 1);
        off.x = or_sampler_next(&s);
        off.y = or_sampler_next(&s);
    }
    *sp = s;
    return camera_ray(sc, ((float)px + off.x) / (float)rs->width, ((float)py + off.y) / (float)rs->height);
}

/* the camera ray of one sample of every pixel, as ymrt_ray records */
void or_primary_rays(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t sample, ymrt_ray *out)
{
    for (uint32_t y = 0; y < rs->height; ++y)
        for (uint32_t x = 0; x < rs->width; ++x) {
            or_sampler s;
            or_ray r = primary_ray(sc, rs, frame, x, y, sample, &s);
            ymrt_ray *o = &out[(size_t)y * rs->width + x];
            o->ox = r.o.x; o->oy = r.o.y; o->oz = r.o.z; o->tmin = OR_TMIN;
            o->dx = r.d.x; o->dy = r.d.y; o->dz = r.d.z; o->tmax = OR_TMAX;
        }
}

/* instance hit first by sample 0's camera ray of every pixel (the picking
 * buffer of RayTraceEditorView, src/Shading.cu:1545-1547) */
void or_first_hits(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t *ids)
{
    for (uint32_t y = 0; y < rs->height; ++y)
        for (uint32_t x = 0; x < rs->width; ++x) {
            or_sampler s;
            or_ray ray = primary_ray(sc, rs, frame, x, y, 0, &s);
            or_hit h;
            or_nearby nb[OR_MAX_NEARBY];
            int nc = 0;
            ids[(size_t)y * rs->width + x] = or_trace_ray(sc, &ray, &h, &nc, nb) ? h.inst : OR_EMPTY;
        }
}

/* A DefaultMtl whose only lobe is Lambert with an untextured albedo: every
 * path term after a vertex on it is linear in that albedo (cosine sampling,
 * one lobe, MIS pdfs without the albedo), so d(term)/d(albedo) = term/albedo
 * per channel. */
int or_lambert_only(const ymrt_material_desc *m)
{
    return m->type == YMRT_MATERIAL_DEFAULT && m->specular_weight == 0.0f && m->metalness == 0.0f &&
           m->transmission_weight == 0.0f && m->coat_weight == 0.0f && m->diffuse_albedo_tex == OR_EMPTY &&
           m->specular_weight_tex == OR_EMPTY && m->metalness_tex == OR_EMPTY &&
           m->transmission_weight_tex == OR_EMPTY && m->coat_weight_tex == OR_EMPTY;
}

/* A Lambert-only material but for a diffuse albedo texture. */
static int lambert_textured(const ymrt_material_desc *m)
{
    return m->type == YMRT_MATERIAL_DEFAULT && m->specular_weight == 0.0f && m->metalness == 0.0f &&
           m->transmission_weight == 0.0f && m->coat_weight == 0.0f && m->diffuse_albedo_tex != OR_EMPTY &&
           m->specular_weight_tex == OR_EMPTY && m->metalness_tex == OR_EMPTY &&
           m->transmission_weight_tex == OR_EMPTY && m->coat_weight_tex == OR_EMPTY;
}

/* Metalness 1 without coat: the BSDF is the GGX conductor lobe alone (the
 * diffuse, specular and transmission lobes carry 1 - metalness); isotropic
 * untextured roughness. */
static int metal_only(const ymrt_material_desc *m)
{
    return m->type == YMRT_MATERIAL_DEFAULT && m->metalness == 1.0f && m->metalness_tex == OR_EMPTY &&
           m->coat_weight == 0.0f && m->coat_weight_tex == OR_EMPTY && m->roughness_x == m->roughness_y &&
           m->roughness_x_tex == OR_EMPTY && m->roughness_y_tex == OR_EMPTY;
}

int or_adjoint_prepare(const or_scene *sc, const uint32_t *alb, uint32_t n_alb, const uint32_t *rough,
                       uint32_t n_rough, uint32_t tex, uint32_t vol, or_adj_targets *tg)
{
    memset(tg, 0, sizeof(*tg));
    uint32_t nm = sc->material_count;
    tg->alb_slot = (int32_t *)malloc(sizeof(int32_t) * (nm ? nm : 1));
    tg->rough_slot = (int32_t *)malloc(sizeof(int32_t) * (nm ? nm : 1));
    for (uint32_t i = 0; i < nm; ++i) tg->alb_slot[i] = tg->rough_slot[i] = -1;
    tg->n_alb = n_alb;
    tg->n_rough = n_rough;
    tg->tex = tex;
    tg->vol = vol;
    int pos = 1;
    for (uint32_t i = 0; i < n_alb; ++i, ++pos) {
        if (alb[i] >= nm || !or_lambert_only(&sc->materials[alb[i]])) return pos;
        tg->alb_slot[alb[i]] = (int32_t)i;
    }
    for (uint32_t i = 0; i < n_rough; ++i, ++pos) {
        if (rough[i] >= nm || !metal_only(&sc->materials[rough[i]])) return pos;
        tg->rough_slot[rough[i]] = (int32_t)i;
    }
    tg->off_rough = (size_t)n_alb * 3;
    tg->off_tex = tg->off_rough + n_rough;
    size_t ntex = 0;
    if (tex != OR_EMPTY) {
        if (tex >= sc->texture_count || sc->textures[tex].desc->type != YMRT_TEXTURE_IMAGE) return pos;
        for (uint32_t m = 0; m < nm; ++m) {
            const ymrt_material_desc *md = &sc->materials[m];
            if (md->diffuse_albedo_tex == tex && !lambert_textured(md)) return pos;
            /* a texture read by any other material slot or a density is not
             * a target either: its gradient would miss those uses */
            const uint32_t other[] = {md->specular_albedo_tex, md->roughness_x_tex, md->roughness_y_tex,
                                      md->specular_weight_tex, md->metalness_tex, md->transmission_weight_tex,
                                      md->coat_albedo_tex, md->coat_weight_tex, md->coat_thickness_tex,
                                      md->coat_roughness_x_tex, md->coat_roughness_y_tex,
                                      md->normal_mapping_tex, md->bump_mapping_tex};
            for (size_t k = 0; k < sizeof(other) / sizeof(other[0]); ++k)
                if (other[k] == tex) return pos;
        }
        for (uint32_t v = 0; v < sc->volume_count; ++v)
            if (sc->volumes[v].density_texture == (int32_t)tex) return pos;
        ntex = (size_t)sc->textures[tex].width * sc->textures[tex].height * sc->textures[tex].channels;
    }
    ++pos;
    tg->off_vox = tg->off_tex + ntex;
    size_t nvox = 0;
    if (vol != OR_EMPTY) {
        if (vol >= sc->volume_count) return pos;
        const ymrt_volume_desc *v = &sc->volumes[vol];
        if (!(v->density_grid >= 0 && (uint32_t)v->density_grid < sc->grid_count) || v->density_texture >= 0 ||
            v->sigma_t[0] != v->sigma_t[1] || v->sigma_t[0] != v->sigma_t[2])
            return pos;
        const or_grid *g = &sc->grids[v->density_grid];
        nvox = (size_t)g->nx * g->ny * g->nz;
    }
    tg->n_params = tg->off_vox + nvox;
    return 0;
}

void or_adjoint_free(or_adj_targets *tg)
{
    free(tg->alb_slot);
    free(tg->rough_slot);
    tg->alb_slot = tg->rough_slot = NULL;
}

/* The four bilinear taps of or_texture_eval at (u, v): texel offsets (of
 * the first channel) and weights. */
static void texture_taps(const or_texture *t, float u, float v, size_t *off, float *w)
{
    float tu = u * t->su + t->ou;
    if (tu < 0.0f) tu = fmaxr(0.0f, tu + fabsf(floorf(tu)));
    float tv = v * t->sv + t->ov;
    if (tv < 0.0f) tv = fmaxr(0.0f, tv + fabsf(floorf(tv)));
    float px = tu * (float)t->width, py = tv * (float)t->height;
    float cx = px - 0.5f, cy = py - 0.5f;
    int x0 = (int)floorf(cx), y0 = (int)floorf(cy);
    float fx = cx - (float)x0, fy = cy - (float)y0;
    int xs[4] = {x0, x0 + 1, x0, x0 + 1}, ys[4] = {y0, y0, y0 + 1, y0 + 1};
    w[0] = (1.0f - fx) * (1.0f - fy); w[1] = fx * (1.0f - fy); w[2] = (1.0f - fx) * fy; w[3] = fx * fy;
    int wd = (int)t->width, ht = (int)t->height;
    for (int k = 0; k < 4; ++k) {
        int x = xs[k], y = ys[k];
        if (t->wrap == YMRT_WRAP_CLAMP) {
            x = x < 0 ? 0 : (x > wd - 1 ? wd - 1 : x);
            y = y < 0 ? 0 : (y > ht - 1 ? ht - 1 : y);
        } else {
            x = ((x % wd) + wd) % wd;
            y = ((y % ht) + ht) % ht;
        }
        off[k] = ((size_t)y * wd + x) * t->channels;
    }
}

static v3 trace_path(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t px, uint32_t py,
                     uint32_t sample, uint64_t *counts, or_adj_path *adj);

v3 or_trace_path(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t px, uint32_t py,
                 uint32_t sample, uint64_t *counts)
{
    return trace_path(sc, rs, frame, px, py, sample, counts, NULL);
}

/* Path-replay adjoint of the frame: with adj the per-pixel derivative of a
 * loss with respect to the noise image (rgb), adds to grad (laid out by
 * or_adjoint_prepare) the derivative of sum_pixels adj . image over the
 * frame's paths, and writes the frame like or_render_rows.  Each path is
 * traced once, recording its gradient events (see or_adj_path); the
 * downstream events are resolved with the path's radiance at its end.
 * Needs Russian roulette and the indirect clamp off. */
void or_render_adjoint(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t row0, uint32_t row1,
                       const float *adj, const or_adj_targets *tg, double *grad, float *out)
{
    float inv_spp = 1.0f / (float)(rs->spp > 1 ? rs->spp : 1);
    or_adj_path *a = (or_adj_path *)calloc(1, sizeof(or_adj_path));
    a->tg = tg;
    a->grad = grad;
    for (uint32_t y = row0; y < row1 && y < rs->height; ++y) {
        for (uint32_t x = 0; x < rs->width; ++x) {
            size_t pix = (size_t)y * rs->width + x;
            float acc[4] = {0, 0, 0, 0};
            for (int smp = 0; smp < rs->spp; ++smp) {
                a->nd = 0;
                a->nray = 0;
                a->adj = mk3(adj[pix * 3 + 0] * inv_spp, adj[pix * 3 + 1] * inv_spp, adj[pix * 3 + 2] * inv_spp);
                v3 L = trace_path(sc, rs, frame, x, y, (uint32_t)smp, NULL, a);
                acc[0] += L.x * inv_spp;
                acc[1] += L.y * inv_spp;
                acc[2] += L.z * inv_spp;
                acc[3] += inv_spp;
                for (uint32_t i = 0; i < a->nd; ++i) {
                    const adj_down *e = &a->down[i];
                    v3 d = mul3(sub3(L, e->mark), e->w);
                    grad[e->param] += (double)a->adj.x * d.x + (double)a->adj.y * d.y + (double)a->adj.z * d.z;
                }
            }
            if (out) {
                float *o = out + 4 * pix;
                o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
            }
        }
    }
    free(a->down);
    free(a);
}

static v3 trace_path(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t px, uint32_t py,
                     uint32_t sample, uint64_t *counts, or_adj_path *adj)
{
    or_sampler s;
    uint32_t pixel_index = py * rs->width + px;
    or_ray ray = primary_ray(sc, rs, frame, px, py, sample, &s);
    rng_ctx rng;
    rng.s = &s;
    rng.hash_key[0] = pixel_index; rng.hash_key[1] = 0; rng.hash_key[2] = sample; rng.hash_key[3] = (uint32_t)frame;
    rng.hash_ctr = 0;
    path_ctx c = {sc, rs, &s, &rng, counts, adj};

    v3 L = s3(0.0f), beta = s3(1.0f);
    int depth = 1, never_scatter = 1;
    or_transfer t = sc->camera_transfer;
    uint32_t vertex = 0;

    for (;;) {
        /* ---- trace stage */
        if (counts) { counts[1]++; if (depth == 1 && never_scatter) counts[0]++; }
        or_hit h;
        memset(&h, 0, sizeof(h));
        h.bary = s3(1.0f); h.t = OR_TMAX; h.inst = OR_EMPTY; h.tri = OR_EMPTY;
        or_nearby nb[OR_MAX_NEARBY];
        int nc = 0;
        or_ray tr_ray = ray;
        int hit = or_trace_ray(sc, &tr_ray, &h, &nc, nb);
        float vw[OR_MAX_BOUNDARY_RECORD];
        int vi[OR_MAX_BOUNDARY_RECORD];
        float gs[OR_MAX_BOUNDARY_RECORD];
        int nv = 0;
        if (rs->enable_volume_scattering) {
            v3 trw = s3(1.0f);
            float dist = 0.0f;
            nv = tr_volumes(&t, vi, sc->volume_count);
            for (int i = 0; i < nv; ++i) gs[i] = sc->volumes[vi[i]].g;
            if (adj) adj->L = L;
            int hv = sample_volume(sc, &ray, nv, vi, hit ? h.t : OR_TMAX, &trw, &dist, vw, &s, adj);
            hit = hit || hv;
            beta = mul3(beta, trw);
            if (hv) { h.volume = 1; h.bary = add3(ray.o, scl3(ray.d, dist)); }
        }
        if (!hit) {
            if (rs->enable_env_light) L = add3(L, light_clamp(rs, mul3(beta, background(ray.d)), depth));
            return L;
        }
        /* ---- shading stage */
        rng.hash_key[1] = ++vertex;
        rng.hash_ctr = 0;
        if (adj) {
            adj->beta = beta;
            adj->rough_param = -1;
            adj->nray = 0;
            adj->recording = adj->tg->vol != OR_EMPTY;
        }
        if (rs->enable_volume_scattering && h.volume) {
            v3 vpos = h.bary;
            if (rs->enable_distant_light)
                L = add3(L, light_clamp(rs, mul3(beta, eval_volume_distant(&c, nv, gs, vw, &ray, &t, vpos)), depth));
            if (rs->enable_shape_light)
                L = add3(L, light_clamp(rs, mul3(beta, eval_volume_light(&c, nv, gs, vw, &ray, &t, vpos)), depth));
            v3 wi = s3(0.0f);
            float pdf = 0.0f, pw = 0.0f;
            float u0 = or_sampler_next(&s);
            float u1 = or_sampler_next(&s);
            if (!mixed_phase_sample(vw, gs, nv, u0, u1, neg3(ray.d), &pw, &wi, &pdf)) return L;
            beta = scl3(beta, pw);
            float rr = fminr(fmaxr(beta.x, fmaxr(beta.y, beta.z)), 1.0f);
            if (depth + 1 > rs->ray_depth) {
                if (rs->enable_russian_roulette && or_sampler_next(&s) < rr) beta = divs3(beta, fmaxr(rr, OR_MIN_COLOR_EPSILON));
                else return L;
            }
            ray = make_ray(vpos, wi, OR_EMPTY, OR_EMPTY);
            never_scatter = 0;
            depth += 1;
            if (depth > OR_MAX_RAY_DEPTH) return L;
            continue;
        }
        v3 p, perr, sn, ng, dpdu, dpdv, p_obj;
        v2 uv;
        fetch_shading(sc, &h, &p, &perr, &sn, &ng, &uv, &dpdu, &dpdv, &p_obj);
        const ymrt_instance_desc *in = &sc->instances[h.inst];
        if (!(in->material < sc->material_count)) return L;
        const ymrt_material_desc *m = &sc->materials[in->material];
        uint32_t prio;
        fetch_ior(m, &prio);
        if ((t.count != 0 && prio < t.rec[t.max_idx].priority) || in->treat_as_boundary) {
            boundary_bunch(sc, &t, nb, nc);
            ray = make_ray(or_offset_ray_origin(p, perr, ray.d, ng), ray.d, h.inst, h.tri);
            continue;
        }
        if (m->type == YMRT_MATERIAL_LIGHT) {
            if (never_scatter)
                L = add3(L, light_clamp(rs, scl3(mul3(beta, mk3(m->light_color[0], m->light_color[1], m->light_color[2])), m->intensity), depth));
            return L;
        }
        mbsdf b;
        memset(&b, 0, sizeof(b));
        float duv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        or_tex_coord tc;
        tc.st = uv; tc.p_world = p; tc.p_object = p_obj;
        if (m->type == YMRT_MATERIAL_DEFAULT && m->normal_mapping_tex == OR_EMPTY && m->bump_mapping_tex != OR_EMPTY)
            camera_uv_differentials(sc, p, sn, ng, dpdu, dpdv, rs->width, rs->height, 0.25f, 0.25f, duv);
        init_shading_space(sc, &b, m->type == YMRT_MATERIAL_DEFAULT ? m : NULL, h.back, ray.d, ng, sn, dpdu, dpdv,
                           in->world_to_object, tc, duv);
        b.is_default = m->type == YMRT_MATERIAL_DEFAULT;
        if (b.is_default) init_default_bsdf(sc, &b, m, tc, ray.d, h.back, tr_ray_ior(&t), tr_ex_ior(&t, h.inst));
        if (adj) {
            /* a Lambert vertex multiplies everything after it by the albedo:
             * d/d albedo_c = (L_path - L_before) / albedo_c per channel */
            int32_t as = adj->tg->alb_slot[in->material];
            if (as >= 0) {
                for (int ch = 0; ch < 3; ++ch) {
                    float al = m->diffuse_albedo[ch];
                    if (!(al > 0.0f)) continue;
                    v3 w = s3(0.0f);
                    setc3(&w, ch, 1.0f / al);
                    adj_add_down(adj, (uint32_t)(as * 3 + ch), w, L);
                }
            }
            /* the same with the albedo a bilinear blend of four texels */
            if (adj->tg->tex != OR_EMPTY && m->diffuse_albedo_tex == adj->tg->tex && lambert_textured(m)) {
                const or_texture *tx = &sc->textures[adj->tg->tex];
                v3 al = or_texture_graph(sc, m->diffuse_albedo_tex, tc);
                size_t off[4];
                float tw[4];
                texture_taps(tx, uv.x, uv.y, off, tw);
                for (int k = 0; k < 4; ++k) {
                    if (!(tw[k] > 0.0f)) continue;
                    if (tx->channels == 1) {
                        v3 w = mk3(al.x > 0.0f ? tw[k] / al.x : 0.0f, al.y > 0.0f ? tw[k] / al.y : 0.0f,
                                   al.z > 0.0f ? tw[k] / al.z : 0.0f);
                        adj_add_down(adj, (uint32_t)(adj->tg->off_tex + off[k]), w, L);
                    } else {
                        for (int ch = 0; ch < 3; ++ch) {
                            float a_ = comp3(al, ch);
                            if (!(a_ > 0.0f)) continue;
                            v3 w = s3(0.0f);
                            setc3(&w, ch, tw[k] / a_);
                            adj_add_down(adj, (uint32_t)(adj->tg->off_tex + off[k] + (size_t)ch), w, L);
                        }
                    }
                }
            }
            int32_t rsl = adj->tg->rough_slot[in->material];
            if (rsl >= 0) {
                adj->rough_param = (int)(adj->tg->off_rough + (size_t)rsl);
                adj->rough_r = m->roughness_x;
            }
        }
        if (rs->enable_distant_light)
            L = add3(L, light_clamp(rs, mul3(beta, eval_distant_light(&c, &b, ray.d, &h, &t, nc, nb, p, perr, ng)), depth));
        if (rs->enable_shape_light)
            L = add3(L, light_clamp(rs, mul3(beta, eval_shape_light(&c, &b, ray.d, &h, &t, nc, nb, p, perr, ng)), depth));
        v3 wo = to_local(&b, neg3(ray.d)), wi = s3(0.0f), bw = s3(0.0f);
        float pdf = 0.0f;
        float u0 = or_sampler_next(&s);
        float u1 = or_sampler_next(&s);
        float u2 = or_sampler_next(&s);
        if (!bsdf_sample(&b, wo, &bw, &wi, &pdf, u0, u1, u2)) return L;
        /* the sampled direction's BSDF value scales everything after this
         * vertex's next-event deposits (pdf held fixed) */
        if (adj && adj->rough_param >= 0)
            adj_add_down(adj, (uint32_t)adj->rough_param, s3(rough_dlog(adj->rough_r, wo, wi)), L);
        beta = mul3(beta, bw);
        float rr = fminr(fmaxr(beta.x, fmaxr(beta.y, beta.z)), 1.0f);
        if (depth + 1 > rs->ray_depth) {
            if (rs->enable_russian_roulette && or_sampler_next(&s) < rr) beta = divs3(beta, fmaxr(rr, OR_MIN_COLOR_EPSILON));
            else return L;
        }
        v3 nd = to_world(&b, wi);
        int transmit = dot3(neg3(ray.d), ng) * dot3(nd, ng) < 0.0f;
        int invalid = (!same_hemi(wo, wi) && !transmit) || (same_hemi(wo, wi) && transmit);
        if (invalid) return L;
        if (transmit) boundary_bunch(sc, &t, nb, nc);
        ray = make_ray(or_offset_ray_origin(p, perr, nd, ng), nd, h.inst, h.tri);
        never_scatter = 0;
        depth += 1;
        if (depth > OR_MAX_RAY_DEPTH) return L;
    }
}

void or_render_rows(const or_scene *sc, const ymrt_render_settings *rs, int frame, uint32_t row0, uint32_t row1,
                    float *out, uint64_t *counts)
{
    float inv_spp = 1.0f / (float)(rs->spp > 1 ? rs->spp : 1);
    for (uint32_t y = row0; y < row1 && y < rs->height; ++y) {
        for (uint32_t x = 0; x < rs->width; ++x) {
            float acc[4] = {0, 0, 0, 0};
            for (int smp = 0; smp < rs->spp; ++smp) {
                v3 L = or_trace_path(sc, rs, frame, x, y, (uint32_t)smp, counts);
                acc[0] += L.x * inv_spp;
                acc[1] += L.y * inv_spp;
                acc[2] += L.z * inv_spp;
                acc[3] += inv_spp;
            }
            float *o = out + 4 * ((size_t)y * rs->width + x);
            o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
            if (counts) counts[4]++;
        }
    }
}

/* src/Shading.cu:2354-2373 */
void or_accumulate(const float *noise, float *accum, uint32_t w, uint32_t h, int frame)
{
    size_t n = (size_t)w * h * 4;
    for (size_t i = 0; i < n; ++i) {
        if (frame == 1) accum[i] = noise[i];
        else accum[i] = (accum[i] * (float)(frame - 1) + noise[i]) / (float)frame;
    }
}

/* src/Shading.cu:2396-2425 */
void or_tonemap(float *rgba, uint32_t w, uint32_t h, float gamma, float exposure)
{
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        float *p = rgba + 4 * i;
        for (int c = 0; c < 3; ++c) {
            float x = p[c] * exposure;
            float y = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
            p[c] = or_powf(y, 1.0f / gamma);
        }
        p[3] = 1.0f;
    }
}

/* ------------------------------------------------- probes for closed-form tests */
float or_probe_ggx_d(float hx, float hy, float hz, float ax, float ay) { return ggx_d(mk3(hx, hy, hz), ax, ay); }
float or_probe_fresnel_dd(float eta, float c) { return fresnel_dd(eta, c); }
void or_probe_fresnel_dc(const float *eta, const float *k, float c, float *out)
{
    v3 r = fresnel_dc(mk3(eta[0], eta[1], eta[2]), mk3(k[0], k[1], k[2]), c);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void or_probe_edge_tint(const float *refl, const float *tint, float *n, float *k)
{
    v3 nn, kk;
    edge_tint(mk3(refl[0], refl[1], refl[2]), mk3(tint[0], tint[1], tint[2]), &nn, &kk);
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    k[0] = kk.x; k[1] = kk.y; k[2] = kk.z;
}
/* SampleTrHomogeneous with one volume and a fixed uniform u: distance and weight */
int or_probe_free_flight(const float *sigma_t, const float *albedo, float tmax, float u, float *dist, float *w)
{
    or_scene sc;
    memset(&sc, 0, sizeof(sc));
    ymrt_volume_desc v;
    memset(&v, 0, sizeof(v));
    for (int i = 0; i < 3; ++i) { v.sigma_t[i] = sigma_t[i]; v.albedo[i] = albedo[i]; }
    v.density_grid = -1;
    sc.volumes = &v;
    sc.volume_count = 1;
    /* a one-value sampler: Halton index 0 of a table whose prime-2 permutation
     * maps digit 0 to 0 returns the clamp 1e-6; instead feed u through PCG-free
     * path by overriding the sampler below */
    or_sampler s;
    memset(&s, 0, sizeof(s));
    s.type = -1;  /* see or_sampler_next: non-Halton, handled below */
    int vi = 0;
    v3 trw;
    float vw[OR_MAX_BOUNDARY_RECORD];
    g_probe_u = u;
    int r = sample_homogeneous(&sc, 1, &vi, tmax, &trw, dist, vw, &s);
    w[0] = trw.x; w[1] = trw.y; w[2] = trw.z;
    return r;
}
/* Camera::TextureCoordinatesDifferential of the scene's camera at one point */
void or_probe_uv_differentials(const or_scene *sc, const float *hp, const float *sn, const float *ng,
                               const float *dpdu, const float *dpdv, uint32_t width, uint32_t height, float sx,
                               float sy, float *out)
{
    camera_uv_differentials(sc, mk3(hp[0], hp[1], hp[2]), mk3(sn[0], sn[1], sn[2]), mk3(ng[0], ng[1], ng[2]),
                            mk3(dpdu[0], dpdu[1], dpdu[2]), mk3(dpdv[0], dpdv[1], dpdv[2]), width, height, sx, sy,
                            out);
}
/* indirect_light_clamp of one contribution (src/Shading.cu:61-81) */
void or_probe_light_clamp(const ymrt_render_settings *rs, const float *c, int depth, float *out)
{
    v3 r = light_clamp(rs, mk3(c[0], c[1], c[2]), depth);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

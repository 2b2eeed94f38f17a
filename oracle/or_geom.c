/*
 * or_geom.c -- oracle restatement of YMRT's BVH build and ray traversal.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP traversal).
 *
 *   - bounding boxes, BoundFix, slab test: src/BoundingBox.h:10-155
 *   - triangle bounds/centres: src/TriangleDefines.h:14-44
 *   - binned SAH (24 bins, cost n*area), Hoare partition, heap median split,
 *     breadth-first build with median fallback, DFS flatten:
 *     src/BottomBVH.h:41-185, src/BottomBVH.cpp:6-166, src/TopBVH.cpp:34-198
 *   - watertight triangle test: src/ShadingUtilities.cuh:54-134
 *   - mesh traversal (32-entry stack, near child first, skip of the previous
 *     hit triangle): src/ShadingUtilities.cuh:136-217
 *   - TraceRay with the +-2 ulp nearby-hit list: src/ShadingUtilities.cuh:310-431
 *   - TraceVolume: src/ShadingUtilities.cuh:434-559
 *   - BVHTraverseShadow: src/ShadingUtilities.cuh:563-635
 *   - TransformRay: src/Ray.h:57-67
 *   - FetchMeshGeometryNormal / OffsetRayOrigin:
 *     src/ShadingUtilities.cuh:894-951, 993-1011
 *
 * Indices are reported as the ORIGINAL (input) triangle and instance indices:
 * the reference reorders both arrays in place during its builds, so its raw
 * indices are a build-order permutation of these.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle.h"

/* ------------------------------------------------------------------ bbox */
static or_bbox bb_empty(void)
{
    or_bbox b;
    b.pmax = s3(OR_FLOAT_MIN);
    b.pmin = s3(OR_FLOAT_MAX);
    return b;
}
static or_bbox bb_extend(or_bbox b, v3 p)
{
    or_bbox r;
    r.pmin = min3(b.pmin, p);
    r.pmax = max3(b.pmax, p);
    return r;
}
static or_bbox bb_union(or_bbox a, or_bbox b)
{
    or_bbox r;
    r.pmin = min3(a.pmin, b.pmin);
    r.pmax = max3(a.pmax, b.pmax);
    return r;
}
static or_bbox bb_fix(or_bbox b)
{
    const float eps = 0.0005f;
    for (int i = 0; i < 3; ++i) {
        float mx = comp3(b.pmax, i), mn = comp3(b.pmin, i);
        if (mx < mn) { float t = mx; mx = mn; mn = t; }
        if (fabsf(mx - mn) < 2.0f * eps) { mx += eps; mn -= eps; }
        setc3(&b.pmax, i, mx);
        setc3(&b.pmin, i, mn);
    }
    return b;
}
static float bb_area(or_bbox b)
{
    v3 e = sub3(b.pmax, b.pmin);
    return 2.f * (e.x * e.y + e.y * e.z + e.x * e.z);
}
static int bb_longest(or_bbox b)
{
    v3 d = sub3(b.pmax, b.pmin);
    int a = 0;
    if (d.y > comp3(d, a)) a = 1;
    if (d.z > comp3(d, a)) a = 2;
    return a;
}
static v3 bb_center(or_bbox b) { return scl3(add3(b.pmax, b.pmin), 0.5f); }

/* src/BoundingBox.h:119-155 */
static int bb_intersect(or_bbox b, const or_ray *r, float *t)
{
    v3 inv = div3(s3(1.f), r->d);
    float g = 1.0f + 2.0f * error_gamma(3);
    float tx0 = (b.pmin.x - r->o.x) * inv.x, tx1 = (b.pmax.x - r->o.x) * inv.x;
    if (tx0 > tx1) { float s = tx0; tx0 = tx1; tx1 = s; tx1 *= g; }
    float ty0 = (b.pmin.y - r->o.y) * inv.y, ty1 = (b.pmax.y - r->o.y) * inv.y;
    if (ty0 > ty1) { float s = ty0; ty0 = ty1; ty1 = s; ty1 *= g; }
    float tz0 = (b.pmin.z - r->o.z) * inv.z, tz1 = (b.pmax.z - r->o.z) * inv.z;
    if (tz0 > tz1) { float s = tz0; tz0 = tz1; tz1 = s; tz1 *= g; }
    float ti = tx0;
    if (ty0 > ti) ti = ty0;
    if (tz0 > ti) ti = tz0;
    float to = tx1;
    if (ty1 < to) to = ty1;
    if (tz1 < to) to = tz1;
    int hit = (to > ti) && (to > 0.0f) && (ti <= r->t);
    *t = hit ? (ti < 0.0f ? to : ti) : OR_TMAX;
    return hit;
}

/* --------------------------------------------------------------- builder */
typedef struct {
    or_bbox bbox;
    v3 center;
    uint32_t idx;
} prim_info;

static float cc(const prim_info *p, int axis) { return comp3(p->center, axis); }

static uint32_t partition_prims(prim_info *p, uint32_t count, float pivot, int axis)
{
    uint32_t left = 0, right = count - 1;
    while (left <= right && right != OR_EMPTY) {
        if (cc(&p[left], axis) <= pivot) {
            ++left;
        } else {
            prim_info t = p[left]; p[left] = p[right]; p[right] = t;
            --right;
        }
    }
    return right;
}

static int eval_sah(const prim_info *p, uint32_t count, or_bbox cb, float parent_cost,
                    float *split_pos, int *split_axis)
{
    enum { BINS = 24 };
    int best_axis = -1;
    float best_pos = 0.0f, best_cost = 1E36f;
    struct { uint32_t count; or_bbox bbox; } bins[BINS][3];
    for (int i = 0; i < BINS; ++i)
        for (int a = 0; a < 3; ++a) { bins[i][a].count = 0; bins[i][a].bbox = bb_empty(); }
    int valid[3] = {1, 1, 1};
    v3 ext = sub3(cb.pmax, cb.pmin);
    for (int a = 0; a < 3; ++a) {
        if (comp3(ext, a) < OR_FLOAT_EPSILON) { valid[a] = 0; setc3(&ext, a, OR_FLOAT_EPSILON); }
    }
    v3 iext = div3(s3(1.0f), ext);
    v3 bin_len = divs3(max3(sub3(cb.pmax, cb.pmin), s3(0.0f)), (float)BINS);
    for (uint32_t i = 0; i < count; ++i) {
        v3 ndc = mul3(sub3(p[i].center, cb.pmin), iext);
        for (int a = 0; a < 3; ++a) {
            int bi = (int)(comp3(ndc, a) * BINS);
            if (bi > BINS - 1) bi = BINS - 1;
            if (bi < 0) bi = 0;
            bins[bi][a].count++;
            bins[bi][a].bbox = bb_union(bins[bi][a].bbox, p[i].bbox);
        }
    }
    float la[BINS - 1][3], ra[BINS - 1][3];
    uint32_t lc[BINS - 1][3], rc[BINS - 1][3];
    for (int a = 0; a < 3; ++a) {
        if (!valid[a]) continue;
        or_bbox lb = bb_empty(), rb = bb_empty();
        uint32_t lcount = 0, rcount = 0;
        for (int i = 0; i < BINS - 1; ++i) {
            lcount += bins[i][a].count;
            lb = bb_union(lb, bins[i][a].bbox);
            lc[i][a] = lcount;
            la[i][a] = bb_area(lb);
            rcount += bins[BINS - 1 - i][a].count;
            rb = bb_union(rb, bins[BINS - 1 - i][a].bbox);
            rc[BINS - 2 - i][a] = rcount;
            ra[BINS - 2 - i][a] = bb_area(rb);
        }
        for (int i = 0; i < BINS - 1; ++i) {
            float cost = (float)lc[i][a] * la[i][a] + (float)rc[i][a] * ra[i][a];
            if (cost <= 0.0f) continue;
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = a;
                best_pos = comp3(cb.pmin, a) + (float)(i + 1) * comp3(bin_len, a);
            }
        }
    }
    if (best_axis == -1 || best_cost >= parent_cost) return 0;
    *split_axis = best_axis;
    *split_pos = best_pos;
    return 1;
}

static void heap_down(prim_info *p, uint32_t i, uint32_t right, int axis)
{
    for (;;) {
        uint32_t l = 2 * i + 1, r = 2 * i + 2;
        if (!(l <= right)) return;
        uint32_t target = l;
        if (r <= right && cc(&p[r], axis) > cc(&p[l], axis)) target = r;
        if (cc(&p[i], axis) > cc(&p[target], axis)) return;
        prim_info t = p[i]; p[i] = p[target]; p[target] = t;
        i = target;
    }
}

static uint32_t find_mid(prim_info *p, uint32_t count, int axis)
{
    uint32_t n_th = (count - 1) / 2;
    uint32_t n_pass = count - n_th;
    for (uint32_t i = count / 2 - 1; i != OR_EMPTY; --i) heap_down(p, i, count - 1, axis);
    for (uint32_t i = count - 1; i > 0 && n_pass > 0; --i, --n_pass) {
        prim_info t = p[0]; p[0] = p[i]; p[i] = t;
        heap_down(p, 0, i - 1, axis);
    }
    return n_th;
}

typedef struct {
    or_bbox bbox;
    uint32_t left, right, offset, count;
} raw_node;

static uint32_t flatten_rec(const raw_node *raw, uint32_t ri, or_bnode *out, uint32_t *n)
{
    uint32_t i = (*n)++;
    out[i].bbox = raw[ri].bbox;
    if (raw[ri].left == OR_EMPTY || raw[ri].right == OR_EMPTY) {
        out[i].a = raw[ri].offset;
        out[i].b = raw[ri].count;
        return i;
    }
    flatten_rec(raw, raw[ri].left, out, n);
    out[i].a = OR_EMPTY;
    out[i].b = flatten_rec(raw, raw[ri].right, out, n);
    return i;
}

/* Builds the reference's binary BVH over prims (reordered in place). */
static or_bnode *build_bvh(prim_info *prims, uint32_t count, uint32_t *node_count)
{
    uint32_t cap = 2 * count - 1;
    raw_node *raw = (raw_node *)calloc(cap, sizeof(raw_node));
    char *mid = (char *)calloc(cap, 1);
    raw[0].offset = 0;
    raw[0].count = count;
    uint32_t total = 1, start = 0;
    while (total - start > 0) {
        uint32_t end = total;
        for (uint32_t ni = start; ni < end; ++ni) {
            raw_node *rn = &raw[ni];
            prim_info *np = prims + rn->offset;
            or_bbox nb = bb_empty(), cb = bb_empty();
            for (uint32_t i = 0; i < rn->count; ++i) {
                nb = bb_union(nb, np[i].bbox);
                cb = bb_extend(cb, np[i].center);
            }
            rn->bbox = nb;
            if (rn->count <= 1) { rn->left = rn->right = OR_EMPTY; continue; }
            int do_mid = mid[ni];
            uint32_t split_index = 0;
            if (!do_mid) {
                float parent_cost = (float)rn->count * bb_area(rn->bbox);
                int axis; float pos;
                if (!eval_sah(np, rn->count, cb, parent_cost, &pos, &axis)) {
                    do_mid = 1;
                } else {
                    split_index = partition_prims(np, rn->count, pos, axis);
                    if (split_index == OR_EMPTY || split_index == rn->count - 1) do_mid = 1;
                }
            }
            uint32_t l = total, r = total + 1;
            total += 2;
            rn->left = l;
            rn->right = r;
            if (do_mid) {
                int axis = bb_longest(nb);
                uint32_t m = find_mid(np, rn->count, axis);
                raw[l].offset = rn->offset; raw[l].count = m + 1;
                raw[r].offset = rn->offset + m + 1; raw[r].count = rn->count - 1 - m;
                mid[l] = 1; mid[r] = 1;
            } else {
                raw[l].offset = rn->offset; raw[l].count = split_index + 1;
                raw[r].offset = rn->offset + split_index + 1; raw[r].count = rn->count - 1 - split_index;
            }
        }
        start = end;
    }
    or_bnode *nodes = (or_bnode *)calloc(total, sizeof(or_bnode));
    uint32_t n = 0;
    flatten_rec(raw, 0, nodes, &n);
    free(raw);
    free(mid);
    *node_count = n;
    return nodes;
}

static v3 ld3(const float *a, uint32_t i) { return mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

static void tri_verts(const or_mesh *m, uint32_t tri, v3 *p0, v3 *p1, v3 *p2)
{
    *p0 = ld3(m->pos, m->pidx[3 * tri]);
    *p1 = ld3(m->pos, m->pidx[3 * tri + 1]);
    *p2 = ld3(m->pos, m->pidx[3 * tri + 2]);
}

void or_shading_init(or_scene *sc, const ymrt_scene_desc *desc); /* or_shade.c */
void or_shading_free(or_scene *sc);

or_scene *or_scene_create(const ymrt_scene_desc *d)
{
    or_scene *sc = (or_scene *)calloc(1, sizeof(or_scene));
    sc->mesh_count = d->mesh_count;
    sc->meshes = (or_mesh *)calloc(d->mesh_count ? d->mesh_count : 1, sizeof(or_mesh));
    for (uint32_t mi = 0; mi < d->mesh_count; ++mi) {
        const ymrt_mesh_desc *md = &d->meshes[mi];
        or_mesh *m = &sc->meshes[mi];
        m->tri_count = md->triangle_count;
        m->pidx = md->position_indices;
        m->pos = md->positions;
        m->nidx = md->normal_indices;
        m->nrm = md->normals;
        m->tidx = md->texcoord_indices;
        m->uv = md->texcoords;
        prim_info *pi = (prim_info *)malloc(sizeof(prim_info) * (m->tri_count ? m->tri_count : 1));
        for (uint32_t t = 0; t < m->tri_count; ++t) {
            v3 p0, p1, p2;
            tri_verts(m, t, &p0, &p1, &p2);
            or_bbox b = bb_empty();
            b = bb_extend(b, p0); b = bb_extend(b, p1); b = bb_extend(b, p2);
            pi[t].bbox = bb_fix(b);
            pi[t].center = divs3(add3(add3(p0, p1), p2), 3.f);
            pi[t].idx = t;
        }
        if (m->tri_count) {
            m->nodes = build_bvh(pi, m->tri_count, &m->node_count);
            m->bounds = m->nodes[0].bbox;
        } else {
            m->bounds = bb_empty();
        }
        m->order = (uint32_t *)malloc(sizeof(uint32_t) * (m->tri_count ? m->tri_count : 1));
        for (uint32_t t = 0; t < m->tri_count; ++t) m->order[t] = pi[t].idx;
        free(pi);
    }
    sc->sphere_count = d->sphere_count;
    sc->spheres = d->spheres;
    sc->instance_count = d->instance_count;
    sc->instances = d->instances;
    if (d->instance_count) {
        prim_info *pi = (prim_info *)malloc(sizeof(prim_info) * d->instance_count);
        for (uint32_t i = 0; i < d->instance_count; ++i) {
            const ymrt_instance_desc *in = &d->instances[i];
            or_bbox ob;
            if (in->geometry_type == YMRT_GEOMETRY_SPHERE) {
                float r = d->spheres[in->geometry].radius;
                ob.pmin = mk3(-r, -r, -r);
                ob.pmax = mk3(r, r, r);
            } else {
                ob = sc->meshes[in->geometry].bounds;
            }
            or_bbox wb = bb_empty();
            for (int c = 0; c < 8; ++c) {
                v3 corner = mk3((c & 1) ? ob.pmax.x : ob.pmin.x, (c & 2) ? ob.pmax.y : ob.pmin.y,
                                (c & 4) ? ob.pmax.z : ob.pmin.z);
                wb = bb_extend(wb, m4_point(in->object_to_world, corner));
            }
            pi[i].bbox = bb_fix(wb);
            pi[i].center = bb_center(pi[i].bbox);
            pi[i].idx = i;
        }
        sc->top = build_bvh(pi, d->instance_count, &sc->top_count);
        sc->inst_order = (uint32_t *)malloc(sizeof(uint32_t) * d->instance_count);
        for (uint32_t i = 0; i < d->instance_count; ++i) sc->inst_order[i] = pi[i].idx;
        free(pi);
    }
    or_shading_init(sc, d);
    return sc;
}

void or_scene_destroy(or_scene *sc)
{
    if (!sc) return;
    for (uint32_t i = 0; i < sc->mesh_count; ++i) {
        free(sc->meshes[i].nodes);
        free(sc->meshes[i].order);
    }
    free(sc->meshes);
    free(sc->top);
    free(sc->inst_order);
    or_shading_free(sc);
    free(sc);
}

/* ------------------------------------------------------------- intersection */
static v3 permute3(v3 p, int x, int y, int z) { return mk3(comp3(p, x), comp3(p, y), comp3(p, z)); }
static int max_dim(v3 v)
{
    int m = 0;
    if (v.y > comp3(v, m)) m = 1;
    if (v.z > comp3(v, m)) m = 2;
    return m;
}

/* src/ShadingUtilities.cuh:54-134.  Returns 1 with t/bary/back filled when the
 * plane crossing lies inside the triangle and denominator != 0 (t unchecked). */
int or_intersect_triangle(const v3 q0, const v3 q1, const v3 q2, const or_ray *ray,
                          float *t_out, v3 *bary, int *back)
{
    v3 gn = normalize3(cross3(sub3(q1, q0), sub3(q2, q0)));
    v3 p0 = sub3(q0, ray->o), p1 = sub3(q1, ray->o), p2 = sub3(q2, ray->o);
    int kz = max_dim(abs3(ray->d));
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    v3 d = permute3(ray->d, kx, ky, kz);
    p0 = permute3(p0, kx, ky, kz);
    p1 = permute3(p1, kx, ky, kz);
    p2 = permute3(p2, kx, ky, kz);
    float sx = -d.x / d.z, sy = -d.y / d.z, sz = 1.f / d.z;
    p0.x += sx * p0.z; p0.y += sy * p0.z;
    p1.x += sx * p1.z; p1.y += sy * p1.z;
    p2.x += sx * p2.z; p2.y += sy * p2.z;
    float e0 = p1.x * p2.y - p1.y * p2.x;
    float e1 = p2.x * p0.y - p2.y * p0.x;
    float e2 = p0.x * p1.y - p0.y * p1.x;
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return 0;
    float den = e0 + e1 + e2;
    if (den == 0) return 0;
    p0.z *= sz; p1.z *= sz; p2.z *= sz;
    float inv = 1.0f / den;
    float t = e0 * p0.z + e1 * p1.z + e2 * p2.z;
    t *= inv;
    *t_out = t;
    *bary = scl3(mk3(e0, e1, e2), inv);
    *back = dot3(gn, ray->d) > 0.0f;
    return 1;
}

typedef struct { float t; int hit; int back; } extra_info;
static int intersect_mesh(const or_mesh *m, or_ray *ray, uint32_t inst, or_hit *hr, int any, extra_info *ex);

/* IntersectSphereObject (src/ShadingUtilities.cuh:219-275), object space.  The
 * roots are divided by 2a in double precision as the reference's 2.0 literal
 * makes them; hr->bary receives the object-space hit point. */
static int intersect_sphere(float radius, or_ray *ray, or_hit *hr, extra_info *ex)
{
    v3 o = ray->o, d = ray->d;
    float a = d.x * d.x + d.y * d.y + d.z * d.z;
    float b = 2.0f * (d.x * o.x + d.y * o.y + d.z * o.z);
    float c = o.x * o.x + o.y * o.y + o.z * o.z - radius * radius;
    float disc = b * b - 4.0f * a * c;
    if (disc < 0) return 0;
    float sq = sqrtf(disc);
    float t1 = (float)((double)(-b - sq) / (2.0 * (double)a));
    float t2 = (float)((double)(-b + sq) / (2.0 * (double)a));
    if (ex) {
        ex->hit = 1;
        ex->t = t1 >= OR_TMIN ? t1 : t2;
        ex->back = dot3(d, normalize3(add3(o, scl3(d, ex->t)))) > 0.0f;
    }
    if (t1 > ray->t || t2 < OR_TMIN) return 0;
    float t = t1;
    if (t < OR_TMIN) {
        t = t2;
        if (t > ray->t) return 0;
    }
    ray->t = t;
    v3 po = add3(o, scl3(d, t));
    float dist = length3(po);
    po = dist > OR_FLOAT_EPSILON ? divs3(scl3(po, radius), dist) : s3(0.0f);
    hr->t = t;
    hr->bary = po;
    hr->back = dot3(d, divs3(po, radius)) > 0.0f;
    hr->tri = OR_EMPTY;
    return 1;
}

int or_sphere_any(float radius, or_ray *ray)
{
    or_hit h;
    return intersect_sphere(radius, ray, &h, NULL);
}

static int intersect_instance(const or_scene *sc, const ymrt_instance_desc *in, uint32_t ii, or_ray *orr,
                              or_hit *ph, int any, extra_info *ex)
{
    if (in->geometry_type == YMRT_GEOMETRY_SPHERE)
        return intersect_sphere(sc->spheres[in->geometry].radius, orr, ph, ex);
    return intersect_mesh(&sc->meshes[in->geometry], orr, ii, ph, any, ex);
}

/* src/ShadingUtilities.cuh:136-217 with IntersectTri folded in. */
static int intersect_mesh(const or_mesh *m, or_ray *ray, uint32_t inst, or_hit *hr,
                          int any, extra_info *ex)
{
    if (!m->nodes) return 0;
    uint32_t stack[64];
    int top = -1;
    int skip = ray->last_inst != OR_EMPTY && ray->last_tri != OR_EMPTY;
    int hit = 0;
    stack[++top] = 0;
    while (top != -1) {
        uint32_t ni = stack[top--];
        const or_bnode *n = &m->nodes[ni];
        if (n->a != OR_EMPTY) {
            for (uint32_t k = 0; k < n->b; ++k) {
                uint32_t tri = m->order[n->a + k];
                if (skip && inst == ray->last_inst && tri == ray->last_tri) continue;
                v3 p0, p1, p2, bary;
                float t; int back;
                tri_verts(m, tri, &p0, &p1, &p2);
                if (!or_intersect_triangle(p0, p1, p2, ray, &t, &bary, &back)) continue;
                if (ex && t >= OR_TMIN) {
                    ex->hit = 1;
                    ex->t = fminr(ex->t, t);
                    ex->back = back;
                }
                if (t < OR_TMIN || t > ray->t) continue;
                ray->t = t;
                hr->t = t;
                hr->bary = bary;
                hr->back = back;
                hr->tri = tri;
                hit = 1;
                if (any) return 1;
            }
            continue;
        }
        uint32_t nn = ni + 1, fn = n->b;
        float tn, tf;
        int in = bb_intersect(m->nodes[nn].bbox, ray, &tn);
        int inf_ = bb_intersect(m->nodes[fn].bbox, ray, &tf);
        if (tf < tn) {
            float s = tf; tf = tn; tn = s;
            uint32_t u = fn; fn = nn; nn = u;
            int b = inf_; inf_ = in; in = b;
        }
        if (inf_) stack[++top] = fn;
        if (in) stack[++top] = nn;
    }
    return hit;
}

static void transform_ray(const or_ray *r, const float *m, or_ray *o)
{
    o->o = m4_point(m, r->o);
    o->d = m4_dir(m, r->d);
    o->t = r->t;
    o->last_inst = r->last_inst;
    o->last_tri = r->last_tri;
}

#define NEARBY_PUSH(lst, cnt, P, T, B) do { if ((cnt) < OR_MAX_NEARBY) { \
    (lst)[(cnt)].prim = (P); (lst)[(cnt)].t = (T); (lst)[(cnt)].back = (B); ++(cnt); } } while (0)

/* mode 0: all instances (TraceRay), mode 1: boundary instances only (TraceVolume) */
static int trace_nearby(const or_scene *sc, or_ray *ray, or_hit *hr, int *ncount, or_nearby *nearby, int mode)
{
    *ncount = 0;
    if (!sc->top) return 0;
    int nc = 0;
    float tub = 0.0f, tlb = 0.0f;
    or_nearby lst[OR_MAX_NEARBY];
    uint32_t stack[64];
    int top = -1, hit = 0;
    stack[++top] = 0;
    while (top != -1) {
        uint32_t ni = stack[top--];
        const or_bnode *n = &sc->top[ni];
        if (n->a != OR_EMPTY) {
            for (uint32_t k = 0; k < n->b; ++k) {
                uint32_t ii = sc->inst_order[n->a + k];
                const ymrt_instance_desc *in = &sc->instances[ii];
                if (mode == 1 && !in->treat_as_boundary) continue;
                or_ray orr;
                transform_ray(ray, in->world_to_object, &orr);
                or_hit ph;
                memset(&ph, 0, sizeof(ph));
                ph.t = OR_TMAX; ph.inst = OR_EMPTY; ph.tri = OR_EMPTY;
                extra_info ex = {OR_TMAX, 0, 0};
                if (intersect_instance(sc, in, ii, &orr, &ph, 0, &ex)) {
                    hr->back = ph.back;
                    hr->bary = ph.bary;
                    hr->t = ph.t;
                    hr->tri = ph.tri;
                    hr->inst = ii;
                    hr->volume = 0;
                    hit = 1;
                    ray->t = orr.t;
                    tub = next_float_up(next_float_up(ray->t));
                    tlb = next_float_down(next_float_down(ray->t));
                    int head = 0, tail = nc - 1;
                    while (head <= tail) {
                        if (lst[head].t > tlb && lst[head].t < tub) {
                            ++head;
                        } else {
                            or_nearby s = lst[head]; lst[head] = lst[tail]; lst[tail] = s;
                            --tail;
                        }
                    }
                    nc = head;
                    NEARBY_PUSH(lst, nc, ii, hr->t, hr->back);
                } else if (hit && ex.hit && ex.t > tlb && ex.t < tub) {
                    NEARBY_PUSH(lst, nc, ii, ex.t, ex.back);
                }
            }
            continue;
        }
        uint32_t nn = ni + 1, fn = n->b;
        float tn, tf;
        int a = bb_intersect(sc->top[nn].bbox, ray, &tn);
        int b = bb_intersect(sc->top[fn].bbox, ray, &tf);
        if (tf < tn) {
            float s = tf; tf = tn; tn = s;
            uint32_t u = fn; fn = nn; nn = u;
            int w = b; b = a; a = w;
        }
        if (b) stack[++top] = fn;
        if (a) stack[++top] = nn;
    }
    for (int i = 0; i < nc; ++i) nearby[i] = lst[i];
    *ncount = nc;
    return hit;
}

int or_trace_ray(const or_scene *sc, or_ray *ray, or_hit *hr, int *ncount, or_nearby *nearby)
{
    return trace_nearby(sc, ray, hr, ncount, nearby, 0);
}

int or_trace_volume(const or_scene *sc, or_ray *ray, or_hit *hr, int *ncount, or_nearby *nearby)
{
    if (sc->volume_count == 0) { *ncount = 0; return 0; }
    return trace_nearby(sc, ray, hr, ncount, nearby, 1);
}

int or_trace_shadow(const or_scene *sc, or_ray *ray)
{
    if (!sc->top) return 0;
    uint32_t stack[64];
    int top = -1;
    stack[++top] = 0;
    while (top != -1) {
        uint32_t ni = stack[top--];
        const or_bnode *n = &sc->top[ni];
        if (n->a != OR_EMPTY) {
            for (uint32_t k = 0; k < n->b; ++k) {
                uint32_t ii = sc->inst_order[n->a + k];
                const ymrt_instance_desc *in = &sc->instances[ii];
                if (in->treat_as_boundary) continue;
                or_ray orr;
                transform_ray(ray, in->world_to_object, &orr);
                or_hit ph;
                memset(&ph, 0, sizeof(ph));
                ph.t = OR_TMAX;
                if (intersect_instance(sc, in, ii, &orr, &ph, 1, NULL)) {
                    ray->t = orr.t;
                    return 1;
                }
            }
            continue;
        }
        uint32_t nn = ni + 1, fn = n->b;
        float tn, tf;
        int a = bb_intersect(sc->top[nn].bbox, ray, &tn);
        int b = bb_intersect(sc->top[fn].bbox, ray, &tf);
        if (tf < tn) {
            float s = tf; tf = tn; tn = s;
            uint32_t u = fn; fn = nn; nn = u;
            int w = b; b = a; a = w;
        }
        if (b) stack[++top] = fn;
        if (a) stack[++top] = nn;
    }
    return 0;
}

/* FetchGeometryNormal, src/ShadingUtilities.cuh:878-951 */
void or_fetch_geometry_normal(const or_scene *sc, const or_hit *h, v3 *p, v3 *perr, v3 *ng)
{
    const ymrt_instance_desc *in = &sc->instances[h->inst];
    if (in->geometry_type == YMRT_GEOMETRY_SPHERE) {
        float radius = fmaxr(sc->spheres[in->geometry].radius, OR_FLOAT_EPSILON);
        *p = h->bary;
        *perr = sclf3(error_gamma(5), abs3(h->bary));
        *ng = divs3(h->bary, radius);
    } else {
        const or_mesh *m = &sc->meshes[in->geometry];
        v3 p0, p1, p2;
        tri_verts(m, h->tri, &p0, &p1, &p2);
        v3 b = h->bary;
        float xs = fabsf(p0.x * b.x) + fabsf(p1.x * b.y) + fabsf(p2.x * b.z);
        float ys = fabsf(p0.y * b.x) + fabsf(p1.y * b.y) + fabsf(p2.y * b.z);
        float zs = fabsf(p0.z * b.x) + fabsf(p1.z * b.y) + fabsf(p2.z * b.z);
        *p = add3(add3(scl3(p0, b.x), scl3(p1, b.y)), scl3(p2, b.z));
        *perr = sclf3(error_gamma(7), mk3(xs, ys, zs));
        *ng = cross3(sub3(p1, p0), sub3(p2, p0));
    }
    *p = transform_position(in->object_to_world, *p, perr);
    *ng = normalize3(transpose_transform_vector(in->world_to_object, *ng));
}

/* src/ShadingUtilities.cuh:993-1011 */
v3 or_offset_ray_origin(v3 p, v3 perr, v3 dir, v3 ng)
{
    float dist = dot3(abs3(ng), perr);
    v3 off = scl3(ng, dist);
    if (dot3(dir, ng) < 0.0f) off = neg3(off);
    v3 po = add3(p, off);
    for (int a = 0; a < 3; ++a) {
        float oa = comp3(off, a);
        if (oa > 0.0f) setc3(&po, a, next_float_up(comp3(po, a)));
        else if (oa < 0.0f) setc3(&po, a, next_float_down(comp3(po, a)));
    }
    return po;
}

/* ------------------------------------------------------------------ batches */
static void ray_from_abi(const ymrt_ray *r, or_ray *o)
{
    o->o = mk3(r->ox, r->oy, r->oz);
    o->d = mk3(r->dx, r->dy, r->dz);
    o->t = r->tmax;
    o->last_inst = OR_EMPTY;
    o->last_tri = OR_EMPTY;
}

void or_trace_closest_batch(const or_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits)
{
    for (uint32_t i = 0; i < n; ++i) {
        or_ray r;
        ray_from_abi(&rays[i], &r);
        or_hit h;
        memset(&h, 0, sizeof(h));
        h.t = OR_TMAX; h.inst = OR_EMPTY; h.tri = OR_EMPTY;
        or_nearby nb[OR_MAX_NEARBY];
        int nc;
        if (or_trace_ray(sc, &r, &h, &nc, nb)) {
            hits[i].t = h.t;
            hits[i].triangle = h.tri;
            hits[i].instance = h.inst;
            hits[i].back = (uint32_t)h.back;
        } else {
            hits[i].t = OR_TMAX;
            hits[i].triangle = OR_EMPTY;
            hits[i].instance = OR_EMPTY;
            hits[i].back = 0;
        }
    }
}

void or_trace_any_batch(const or_scene *sc, const ymrt_ray *rays, uint32_t n, uint32_t *occ)
{
    for (uint32_t i = 0; i < n; ++i) {
        or_ray r;
        ray_from_abi(&rays[i], &r);
        occ[i] = (uint32_t)or_trace_shadow(sc, &r);
    }
}

/* Closest hit by testing every triangle of every instance (no BVH); ties keep
 * the first in (instance, triangle) order. */
void or_brute_closest_batch(const or_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits)
{
    for (uint32_t i = 0; i < n; ++i) {
        or_ray wr;
        ray_from_abi(&rays[i], &wr);
        float best = wr.t;
        uint32_t bi = OR_EMPTY, bt = OR_EMPTY, bb = 0;
        for (uint32_t ii = 0; ii < sc->instance_count; ++ii) {
            const ymrt_instance_desc *in = &sc->instances[ii];
            or_ray orr;
            transform_ray(&wr, in->world_to_object, &orr);
            if (in->geometry_type == YMRT_GEOMETRY_SPHERE) {
                or_hit sh;
                orr.t = best;
                if (intersect_sphere(sc->spheres[in->geometry].radius, &orr, &sh, NULL) &&
                    !(sh.t == best && bi != OR_EMPTY)) {
                    best = sh.t; bi = ii; bt = OR_EMPTY; bb = (uint32_t)sh.back;
                }
                continue;
            }
            const or_mesh *m = &sc->meshes[in->geometry];
            for (uint32_t t = 0; t < m->tri_count; ++t) {
                v3 p0, p1, p2, bary;
                float th; int back;
                tri_verts(m, t, &p0, &p1, &p2);
                if (!or_intersect_triangle(p0, p1, p2, &orr, &th, &bary, &back)) continue;
                if (th < OR_TMIN || th > best) continue;
                if (th == best && bi != OR_EMPTY) continue;
                best = th; bi = ii; bt = t; bb = (uint32_t)back;
            }
        }
        hits[i].t = bi == OR_EMPTY ? OR_TMAX : best;
        hits[i].triangle = bt;
        hits[i].instance = bi;
        hits[i].back = bb;
    }
}

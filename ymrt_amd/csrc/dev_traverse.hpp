// dev_traverse.hpp -- two-level wide-BVH traversal for gfx950.
//
// Restates the queries of src/ShadingUtilities.cuh:
//   closest hit  TraceRay            (310-431) incl. the +-2 ulp nearby-hit list
//   any hit      BVHTraverseShadow   (563-635), boundary instances skipped
//   boundaries   TraceVolume         (434-559), only boundary instances
// over a 4-wide tree (one 128-B node per visit, children visited nearest
// first, far children pushed on a per-lane stack in LDS laid out
// [slot][lane] so the 64 lanes of a wave hit 64 distinct banks).  The
// triangle test is the reference's watertight shear test
// (src/ShadingUtilities.cuh:54-134) evaluated in the same operation order with
// contraction off, so t and barycentrics are the reference's floats.  Box
// tests use (b - o) * (1/d) like src/BoundingBox.h:119-155 and scale every
// far distance by 1 + 2*gamma(3), which keeps them conservative.
#pragma once

#include "dev_scene.hpp"

namespace ym {

enum : int { kModeAll = 0, kModeSkipBoundary = 1, kModeBoundaryOnly = 2 };
constexpr int kNearbyCap = 4;

struct RayPre {
    f3 o, d, inv;
    int kx, ky, kz;
    float sx, sy, sz;
};

YM_HD RayPre make_pre(f3 o, f3 d)
{
    RayPre r;
    r.o = o;
    r.d = d;
    r.inv = f3{1.f / d.x, 1.f / d.y, 1.f / d.z};
    f3 a = abs3(d);
    int kz = 0;
    if (a.y > comp(a, kz)) kz = 1;
    if (a.z > comp(a, kz)) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    float dx = comp(d, kx), dy = comp(d, ky), dz = comp(d, kz);
    r.sx = -dx / dz;
    r.sy = -dy / dz;
    r.sz = 1.f / dz;
    return r;
}

struct TriHit {
    float t, e0, e1, e2, inv;
};

// Returns true when the sheared plane crossing lies inside the triangle and
// the denominator is non-zero; t is not range-checked.
YM_HD bool tri_test(const RayPre &r, f3 q0, f3 q1, f3 q2, TriHit &h)
{
    f3 p0 = q0 - r.o, p1 = q1 - r.o, p2 = q2 - r.o;
    f3 a{comp(p0, r.kx), comp(p0, r.ky), comp(p0, r.kz)};
    f3 b{comp(p1, r.kx), comp(p1, r.ky), comp(p1, r.kz)};
    f3 c{comp(p2, r.kx), comp(p2, r.ky), comp(p2, r.kz)};
    a.x += r.sx * a.z;
    a.y += r.sy * a.z;
    b.x += r.sx * b.z;
    b.y += r.sy * b.z;
    c.x += r.sx * c.z;
    c.y += r.sy * c.z;
    float e0 = b.x * c.y - b.y * c.x;
    float e1 = c.x * a.y - c.y * a.x;
    float e2 = a.x * b.y - a.y * b.x;
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float den = e0 + e1 + e2;
    if (den == 0) return false;
    a.z *= r.sz;
    b.z *= r.sz;
    c.z *= r.sz;
    float inv = 1.0f / den;
    float t = e0 * a.z + e1 * b.z + e2 * c.z;
    t *= inv;
    h.t = t;
    h.e0 = e0;
    h.e1 = e1;
    h.e2 = e2;
    h.inv = inv;
    return true;
}

YM_HD bool tri_back(f3 q0, f3 q1, f3 q2, f3 d)
{
    f3 gn = normalize(cross(q1 - q0, q2 - q0));
    return dot(gn, d) > 0.0f;
}

// IntersectSphereObject (src/ShadingUtilities.cuh:219-275) in object space.
// The reference divides the roots by 2.0 * a in double precision and rounds
// to float; for a quotient of two floats that equals the correctly rounded
// float quotient (53 >= 2 * 24 + 2 bits), which is what float division gives
// here (2a is exact).  ex_* describe the first root at or beyond tmin whether
// or not it is accepted (the nearby-hit bookkeeping needs it); on acceptance
// t, the hit point on the sphere and the back-face flag are returned.
struct SphereTest {
    bool ex_hit;
    float ex_t;
    int ex_back;
    float t;
    f3 p;
    int back;
};

YM_HD bool sphere_test(f3 o, f3 d, float radius, float tmin, float tmax, SphereTest &r)
{
    r.ex_hit = false;
    const float a = d.x * d.x + d.y * d.y + d.z * d.z;
    const float b = 2.0f * (d.x * o.x + d.y * o.y + d.z * o.z);
    const float c = o.x * o.x + o.y * o.y + o.z * o.z - radius * radius;
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0) return false;
    const float sq = sqrtf(disc);
    const float t1 = (-b - sq) / (2.0f * a);
    const float t2 = (-b + sq) / (2.0f * a);
    r.ex_hit = true;
    r.ex_t = t1 >= tmin ? t1 : t2;
    r.ex_back = dot(d, normalize(o + d * r.ex_t)) > 0.0f;
    if (t1 > tmax || t2 < tmin) return false;
    float t = t1;
    if (t < tmin) {
        t = t2;
        if (t > tmax) return false;
    }
    f3 po = o + d * t;
    const float dist = length(po);
    po = dist > kFloatEpsilon ? (po * radius) / dist : s3(0.0f);
    r.t = t;
    r.p = po;
    r.back = dot(d, po / radius) > 0.0f;
    return true;
}

// Length of column c of an affine map's linear part: the scale an instance
// applies along object axis c.
YM_HD float sphere_axis_scale(const Affine &m, int c)
{
    return sqrtf(m.r[c] * m.r[c] + m.r[4 + c] * m.r[4 + c] + m.r[8 + c] * m.r[8 + c]);
}

// FetchGeometryNormal (src/ShadingUtilities.cuh:878-951): world position with
// its rounding-error bound and the normalised geometric normal of a hit.  For
// a sphere (b0, b1, b2) is the object-space hit point.
YM_HD void hit_geometry(const DevScene &sc, uint32_t inst, uint32_t tri, float b0, float b1, float b2, f3 &p,
                        f3 &perr, f3 &ng)
{
    const DevInstance &in = sc.instances[inst];
    if (in.flags & kInstSphere) {
        const float radius = gmax(in.radius, kFloatEpsilon);
        p = f3{b0, b1, b2};
        perr = error_gamma(5) * abs3(p);
        ng = p / radius;
    } else {
        const ShadeTri &st = sc.shade[in.shade_base + tri];
        const f3 p0{st.p[0], st.p[1], st.p[2]}, p1{st.p[3], st.p[4], st.p[5]}, p2{st.p[6], st.p[7], st.p[8]};
        const float xs = fabsf(p0.x * b0) + fabsf(p1.x * b1) + fabsf(p2.x * b2);
        const float ys = fabsf(p0.y * b0) + fabsf(p1.y * b1) + fabsf(p2.y * b2);
        const float zs = fabsf(p0.z * b0) + fabsf(p1.z * b1) + fabsf(p2.z * b2);
        p = p0 * b0 + p1 * b1 + p2 * b2;
        perr = error_gamma(7) * f3{xs, ys, zs};
        ng = cross(p1 - p0, p2 - p0);
    }
    p = xf_position(in.o2w, p, &perr);
    ng = normalize(xf_transpose_vector(in.w2o, ng));
}

struct Query {
    f3 o, d;
    float tmin, tmax;
    uint32_t last_inst, last_tri;
};

struct Hit {
    float t;
    uint32_t inst, tri;
    float b0, b1, b2;
    int back;
};

// Per-lane scratch for one traversal: the stack and the nearby list, in LDS
// laid out [slot][lane].  With kWindow > 0 only the top kWindow stack entries
// live in LDS (a ring indexed by depth); an entry pushed deeper than that
// evicts the entry kWindow below it to a global spill area laid out
// [depth - kWindow][thread], and popping restores it.  Most traversals never
// leave the window, so the spill area costs nothing but its allocation while
// the small window lets more waves share a CU's LDS.
//
// With kPreSlots, the world-space ray setup (reciprocal direction, shear
// constants and axis) is parked in LDS [kPreSlots][kBlock] when a traversal
// enters an instance and read back when it leaves, instead of being
// recomputed with six divisions.
constexpr int kPreSlots = 7;
template <int kBlock, int kWindow = 0, bool kPreLds = false>
struct LaneScratch {
    static constexpr int kLanes = kBlock;
    static constexpr bool kParkPre = kPreLds;
    uint32_t *stack;        // [kWindow or kStack][kBlock]
    uint32_t *nb_prim;      // [kNearbyCap][kBlock]
    float *nb_t;
    int *nb_back;
    int lane;
    uint32_t *spill;        // global [kMaxSpill][spill_stride], unused when kWindow == 0
    uint32_t spill_stride, spill_index;
    float *pre_park = nullptr;  // LDS [kPreSlots][kBlock] when kPreLds

    YM_HD void park_pre(const RayPre &r)
    {
        float *q = pre_park + lane;
        q[0 * kBlock] = r.inv.x;
        q[1 * kBlock] = r.inv.y;
        q[2 * kBlock] = r.inv.z;
        q[3 * kBlock] = r.sx;
        q[4 * kBlock] = r.sy;
        q[5 * kBlock] = r.sz;
        q[6 * kBlock] = __builtin_bit_cast(float, r.kz);
    }
    // the parked setup of the ray (o, d)
    YM_HD RayPre unpark_pre(f3 o, f3 d) const
    {
        const float *q = pre_park + lane;
        RayPre r;
        r.o = o;
        r.d = d;
        r.inv = f3{q[0 * kBlock], q[1 * kBlock], q[2 * kBlock]};
        r.sx = q[3 * kBlock];
        r.sy = q[4 * kBlock];
        r.sz = q[5 * kBlock];
        r.kz = __builtin_bit_cast(int, q[6 * kBlock]);
        r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
        r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
        return r;
    }

    static constexpr int kWindowSize = kWindow;
    YM_HD int nbi(int i) const { return i * kBlock + lane; }
    // Three entries at depths sp, sp+1, sp+2 with no spill bookkeeping: only
    // for sp + 3 <= kWindow (or the full stack when kWindow == 0), where those
    // depths are free slots of their own.
    YM_HD void put3(int sp, uint32_t a, uint32_t b, uint32_t c)
    {
        stack[sp * kBlock + lane] = a;
        stack[(sp + 1) * kBlock + lane] = b;
        stack[(sp + 2) * kBlock + lane] = c;
    }
    YM_HD void push(int sp, uint32_t v)
    {
        if constexpr (kWindow == 0) {
            stack[sp * kBlock + lane] = v;
        } else {
            static_assert((kWindow & (kWindow - 1)) == 0, "window must be a power of two");
            const int w = (sp & (kWindow - 1)) * kBlock + lane;
            if (sp >= kWindow) gstore(spill + (size_t)(sp - kWindow) * spill_stride + spill_index, stack[w]);
            stack[w] = v;
        }
    }
    // sp is the depth of the entry being popped (already decremented)
    YM_HD uint32_t pop(int sp)
    {
        if constexpr (kWindow == 0) {
            return stack[sp * kBlock + lane];
        } else {
            const int w = (sp & (kWindow - 1)) * kBlock + lane;
            const uint32_t v = stack[w];
            if (sp >= kWindow) stack[w] = gload(spill + (size_t)(sp - kWindow) * spill_stride + spill_index);
            return v;
        }
    }
};

struct TraceCounters {
    uint32_t nodes, tris;
};

// compare-exchange as selects (no branch, no exec-mask changes)
#define YM_SORT2(ka, va, kb, vb)                      \
    do {                                              \
        const bool sw_ = kb < ka;                     \
        const float lo_ = sw_ ? kb : ka;              \
        const float hi_ = sw_ ? ka : kb;              \
        const uint32_t vlo_ = sw_ ? vb : va;          \
        const uint32_t vhi_ = sw_ ? va : vb;          \
        ka = lo_;                                     \
        kb = hi_;                                     \
        va = vlo_;                                    \
        vb = vhi_;                                    \
    } while (0)

// Resumable traversal: init() then step() until it returns something other
// than kTravContinue.  One step visits one wide node, one top-level leaf
// (an instance entry or a sphere) or one bottom-level leaf, then pops the
// next entry, so a caller can interleave rays (a lane that finishes can
// start another ray while the rest of its wave traverses).
// kSphere: the scene has sphere instances (their test is compiled in only
// then; it costs registers in the traversal loop).
constexpr int kTravContinue = 0, kTravHit = 1, kTravMiss = 2, kTravOverflow = -1;

template <int kMode, bool kAny, bool kNearby, bool kCount, bool kSphere, int kStack>
struct Traversal {
    Query q;
    RayPre pre;
    Hit hit;
    float tbest;
    uint32_t node, cur_inst;
    int sp, nb_count, nbn, ex_back;
    bool found, in_blas, skip_enabled, inst_hit, ex_hit;
    float ex_t, tub, tlb;

    YM_HD int init(const DevScene &sc, const Query &query)
    {
        const uint32_t root = kMode == kModeAll ? sc.tlas_root
                              : kMode == kModeSkipBoundary ? sc.tlas_root_solid
                                                           : sc.tlas_root_boundary;
        q = query;
        found = false;
        if (sc.instance_count == 0 || root == kSentinel) return kTravMiss;
        skip_enabled = q.last_inst != kEmpty && q.last_tri != kEmpty;
        pre = make_pre(q.o, q.d);
        tbest = q.tmax;
        in_blas = false;
        cur_inst = kEmpty;
        inst_hit = false;
        ex_hit = false;
        ex_t = kTMax;
        ex_back = 0;
        tub = 0.0f;
        tlb = 0.0f;
        nbn = 0;
        sp = 0;
        node = root;
        return kTravContinue;
    }

    // nearby-list update when the traversal leaves an instance (TraceRay's
    // per-instance block, src/ShadingUtilities.cuh:357-393)
    template <typename LS>
    YM_HD void nearby_exit(LS &ls)
    {
        if (inst_hit) {
            tub = next_float_up(next_float_up(tbest));
            tlb = next_float_down(next_float_down(tbest));
            int w = 0;
            for (int i = 0; i < nbn; ++i) {
                const float ti = ls.nb_t[ls.nbi(i)];
                if (ti > tlb && ti < tub) {
                    ls.nb_t[ls.nbi(w)] = ti;
                    ls.nb_prim[ls.nbi(w)] = ls.nb_prim[ls.nbi(i)];
                    ls.nb_back[ls.nbi(w)] = ls.nb_back[ls.nbi(i)];
                    ++w;
                }
            }
            nbn = w;
            if (nbn < kNearbyCap) {
                ls.nb_prim[ls.nbi(nbn)] = cur_inst;
                ls.nb_t[ls.nbi(nbn)] = tbest;
                ls.nb_back[ls.nbi(nbn)] = hit.back;
                ++nbn;
            }
        } else if (found && ex_hit && ex_t > tlb && ex_t < tub) {
            if (nbn < kNearbyCap) {
                ls.nb_prim[ls.nbi(nbn)] = cur_inst;
                ls.nb_t[ls.nbi(nbn)] = ex_t;
                ls.nb_back[ls.nbi(nbn)] = ex_back;
                ++nbn;
            }
        }
        nb_count = nbn;
    }

    // leaving an instance: back to world space
    template <typename LS>
    YM_HD void leave_instance(LS &ls)
    {
        if constexpr (LS::kParkPre)
            pre = ls.unpark_pre(q.o, q.d);
        else
            pre = make_pre(q.o, q.d);
        in_blas = false;
        if (kNearby) nearby_exit(ls);
    }

    // Pops the next entry.  kThrough: a sentinel is handled on the spot
    // (leave the instance, pop again); otherwise it is left in node for the
    // next step.
    template <bool kThrough, typename LS>
    YM_HD int pop(LS &ls)
    {
        for (;;) {
            if (sp == 0) return found ? kTravHit : kTravMiss;
            node = ls.pop(--sp);
            if (!kThrough || node != kSentinel) return kTravContinue;
            leave_instance(ls);
        }
    }

    // One wide-node visit (node is an inner node): pushes the far children
    // that are hit and moves to the nearest, or pops.  Without kThrough it
    // never changes the ray setup, so a loop of visits keeps it in place.
    template <bool kThrough, typename LS>
    YM_HD int visit_node(const DevScene &sc, LS &ls, TraceCounters &cnt)
    {
        {
            const WNode *n = sc.nodes + node;
            const float4 lx = gload(reinterpret_cast<const float4 *>(n->lo_x));
            const float4 hx = gload(reinterpret_cast<const float4 *>(n->hi_x));
            const float4 ly = gload(reinterpret_cast<const float4 *>(n->lo_y));
            const float4 hy = gload(reinterpret_cast<const float4 *>(n->hi_y));
            const float4 lz = gload(reinterpret_cast<const float4 *>(n->lo_z));
            const float4 hz = gload(reinterpret_cast<const float4 *>(n->hi_z));
            const uint4 ch = gload(reinterpret_cast<const uint4 *>(n->child));
            if (kCount) cnt.nodes++;
            constexpr float kRobust = 1.0f + 2.0f * error_gamma(3);
            float k0, k1, k2, k3;
#define YM_SLAB(K, LX, HX, LY, HY, LZ, HZ, C)                                   \
    {                                                                           \
        float a0 = (LX - pre.o.x) * pre.inv.x, a1 = (HX - pre.o.x) * pre.inv.x; \
        float b0 = (LY - pre.o.y) * pre.inv.y, b1 = (HY - pre.o.y) * pre.inv.y; \
        float c0 = (LZ - pre.o.z) * pre.inv.z, c1 = (HZ - pre.o.z) * pre.inv.z; \
        float tn = fmaxf(fmaxf(fminf(a0, a1), fminf(b0, b1)), fminf(c0, c1));   \
        float tf = fminf(fminf(fmaxf(a0, a1), fmaxf(b0, b1)), fmaxf(c0, c1));   \
        tf *= kRobust;                                                          \
        bool h = (C != kSentinel) & (tf > tn) & (tf > 0.0f) & (tn <= tbest);    \
        K = h ? (tn < 0.0f ? 0.0f : tn) : INFINITY;                             \
    }
            YM_SLAB(k0, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ch.x)
            YM_SLAB(k1, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ch.y)
            YM_SLAB(k2, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ch.z)
            YM_SLAB(k3, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ch.w)
#undef YM_SLAB
            uint32_t v0 = ch.x, v1 = ch.y, v2 = ch.z, v3 = ch.w;
            YM_SORT2(k0, v0, k1, v1);
            YM_SORT2(k2, v2, k3, v3);
            YM_SORT2(k0, v0, k2, v2);
            YM_SORT2(k1, v1, k3, v3);
            YM_SORT2(k1, v1, k2, v2);
            if (k0 != INFINITY) {
                const int need = (k1 != INFINITY) + (k2 != INFINITY) + (k3 != INFINITY);
                if (sp + need > kStack) return kTravOverflow;
                constexpr int kFast = LS::kWindowSize > 0 ? LS::kWindowSize : kStack;
                if (sp + 3 <= kFast) {
                    // missed children sort last: the hit ones go far to near,
                    // the slots past them are written but stay above the top
                    const uint32_t s0 = need == 3 ? v3 : (need == 2 ? v2 : v1);
                    const uint32_t s1 = need == 3 ? v2 : v1;
                    ls.put3(sp, s0, s1, v1);
                    sp += need;
                } else {
                    if (k3 != INFINITY) ls.push(sp++, v3);
                    if (k2 != INFINITY) ls.push(sp++, v2);
                    if (k1 != INFINITY) ls.push(sp++, v1);
                }
                node = v0;
                return kTravContinue;
            }
        }
        return pop<kThrough>(ls);
    }

    template <typename LS>
    YM_HD int step(const DevScene &sc, LS &ls, TraceCounters &cnt)
    {
        if (node == kSentinel) {
            leave_instance(ls);
            return pop<true>(ls);
        }
        if (!(node & kLeafBit)) {
            return visit_node<true>(sc, ls, cnt);
        } else if (!in_blas) {
            // top-level leaf: one instance
            const uint32_t ii = leaf_first(node);
            const DevInstance in = gload(sc.instances + ii);
            bool take = true;
            if (kMode == kModeSkipBoundary) take = !(in.flags & kInstBoundary);
            if (kMode == kModeBoundaryOnly) take = (in.flags & kInstBoundary) != 0;
            if (kSphere && take && (in.flags & kInstSphere)) {
                SphereTest st;
                const bool acc = sphere_test(xf_point(in.w2o, q.o), xf_dir(in.w2o, q.d), in.radius, q.tmin, tbest, st);
                if (kCount) cnt.tris++;
                if (acc) {
                    tbest = st.t;
                    found = true;
                    hit.t = st.t;
                    hit.inst = ii;
                    hit.tri = kEmpty;
                    if (!kAny) {
                        hit.b0 = st.p.x;
                        hit.b1 = st.p.y;
                        hit.b2 = st.p.z;
                        hit.back = st.back;
                    }
                    if (kAny) return kTravHit;
                }
                if (kNearby) {
                    cur_inst = ii;
                    inst_hit = acc;
                    ex_hit = st.ex_hit;
                    ex_t = st.ex_t;
                    ex_back = st.ex_back;
                    nearby_exit(ls);
                }
            } else if (take) {
                if (sp + 1 > kStack) return kTravOverflow;
                ls.push(sp++, kSentinel);
                if constexpr (LS::kParkPre) ls.park_pre(pre);
                pre = make_pre(xf_point(in.w2o, q.o), xf_dir(in.w2o, q.d));
                in_blas = true;
                cur_inst = ii;
                inst_hit = false;
                ex_hit = false;
                ex_t = kTMax;
                node = in.blas_root;
                return kTravContinue;
            }
        } else {
            const uint32_t first = leaf_first(node), count = leaf_count(node);
            for (uint32_t k = 0; k < count; ++k) {
                const float4 *tp = sc.tris + 3u * (first + k);
                const float4 w0 = gload(tp), w1 = gload(tp + 1), w2 = gload(tp + 2);
                const uint32_t tri = f2u(w0.w);
                if (kCount) cnt.tris++;
                if (skip_enabled && cur_inst == q.last_inst && tri == q.last_tri) continue;
                const f3 q0{w0.x, w0.y, w0.z}, q1{w1.x, w1.y, w1.z}, q2{w2.x, w2.y, w2.z};
                TriHit th;
                if (!tri_test(pre, q0, q1, q2, th)) continue;
                if (kNearby && th.t >= q.tmin) {
                    ex_hit = true;
                    ex_t = gmin(ex_t, th.t);
                    ex_back = tri_back(q0, q1, q2, pre.d);
                }
                if (th.t < q.tmin || th.t > tbest) continue;
                tbest = th.t;
                found = true;
                hit.t = th.t;
                hit.inst = cur_inst;
                hit.tri = tri;
                if (!kAny) {
                    hit.b0 = th.e0 * th.inv;
                    hit.b1 = th.e1 * th.inv;
                    hit.b2 = th.e2 * th.inv;
                    hit.back = tri_back(q0, q1, q2, pre.d);
                }
                inst_hit = true;
                if (kAny) return kTravHit;
            }
        }
        return pop<true>(ls);
    }
};

// Traversal steps of the wave's continuing lanes (r == kTravContinue).
// Node visits and leaf tests are two long code paths; while at least
// kInnerFirst lanes are at wide nodes, only those lanes step, so each path
// runs with a fuller wave instead of both running under partial masks; with
// fewer at nodes every lane steps, so no lane waits for long.  kLoop: the
// node visits run as a loop of their own until fewer than kInnerFirst lanes
// are at nodes (the ray setup then stays in place, which saves register
// copies per visit); faster for closest-hit rays, slower for shadow rays,
// which end at their first hit.  Measured best on the interior frame:
// kInnerFirst 16 with the loop (closest hit), 24 without (shadow rays); 8
// and 32 were slower for both.
template <int kInnerFirst, bool kLoop, class TV, class LS>
__device__ __forceinline__ void wave_step(TV &tv, int &r, const DevScene &sc, LS &ls, TraceCounters &tc)
{
    if constexpr (kLoop) {
        for (;;) {
            const bool at_node = r == kTravContinue && !(tv.node & kLeafBit);
            if (__popcll(__ballot(at_node)) < kInnerFirst) break;
            if (at_node) r = tv.template visit_node<false>(sc, ls, tc);
        }
        if (r == kTravContinue) r = tv.step(sc, ls, tc);
    } else {
        const bool leaf = (tv.node & kLeafBit) != 0;
        const uint64_t inner = __ballot(r == kTravContinue && !leaf);
        if (r == kTravContinue && (!leaf || __popcll(inner) < kInnerFirst)) r = tv.step(sc, ls, tc);
    }
}

// Returns 1 on hit, 0 on miss, -1 on stack overflow (the ray is abandoned).
template <int kMode, bool kAny, bool kNearby, bool kCount, bool kSphere, int kStack, typename LS>
YM_HD int traverse(const DevScene &sc, const Query &q, Hit &hit, int &nb_count, LS &ls, TraceCounters &cnt)
{
    Traversal<kMode, kAny, kNearby, kCount, kSphere, kStack> tv;
    tv.hit = hit;
    tv.nb_count = nb_count;
    int r = tv.init(sc, q);
    while (r == kTravContinue) r = tv.step(sc, ls, cnt);
    hit = tv.hit;
    nb_count = tv.nb_count;
    return r == kTravHit ? 1 : (r == kTravMiss ? 0 : -1);
}

#undef YM_SORT2

}  // namespace ym

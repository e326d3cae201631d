// kernels.hip — gfx950 kernels of libbbgpu.so and the C-ABI entry points that
// launch them: batch traversal (bbgpu_trace) and the wavefront path tracer
// (bbgpu_render).
//
// Traversal follows BVH::RayCast (include/bulbit/bvh.h:101-155) exactly: LIFO
// stack seeded with the root, box test when a node is taken with the current
// t_max, leaf primitives in stored order, far child pushed and near child
// taken next (near = child2 when the ray direction along the split axis is
// negative).  Taking the near child directly instead of pushing and popping it
// tests the same nodes in the same order with the same t_max.  The stack lives
// in LDS, laid out [depth][thread] so that the 64 lanes of a wave touch 64
// consecutive words.
//
// The path tracer is a wavefront pipeline per pass of samples:
//   raygen -> for depth in 0..max_bounces:
//                trace_closest -> shade (emission, NEE, BSDF sample, RR,
//                compaction of continuing paths and shadow rays) -> trace_shadow
//          -> film (per pixel, samples in order)
// Each path keeps its own PCG32 stream seeded as IndependentSampler does, and
// draws in the reference's order, so every sample equals the reference's
// sample up to transcendental rounding.
#include "bb_shade.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace bb
{
void set_error(const char* fmt, ...);
}

namespace bb
{

constexpr int kTraceBlock = 128;
// defaults measured on the 1080p 64-spp frame (one pass per frame): refill a
// wave once 32 of its 64 lanes are idle (from its reservoir: no atomic); keep
// 12 stack entries per lane in LDS (higher occupancy beats fewer overflow
// accesses)
constexpr int kTraceRefillDefault = 32;
constexpr int kShadowRefillDefault = 32;
constexpr int kLdsEntriesDefault = 12;
constexpr int kShadeBlock = 256;
#ifndef BB_RESERVOIR
#define BB_RESERVOIR 128
#endif
constexpr int kReservoir = BB_RESERVOIR; // queries a traversal wave claims per atomic
constexpr int kLdsCdfEntries = 2049; // lights whose CDF the shading kernel keeps in LDS (up to 2048)
constexpr int kMaxGridBlocks = 256 * 32;

// ----------------------------------------------------------------- helpers

BB_D f3 xyz(float4 v) { return mk3(v.x, v.y, v.z); }

// AABB3::TestRay with precomputed inverse direction (bounding_box.h:443-484),
// split into the part that does not depend on t_max and the box's entry
// distance.
//
// The reference tests x, then y against the x-tightened interval, then z
// against the x,y-tightened interval, and ignores any NaN slab distance
// (every comparison with NaN is false).  Unrolled, it fails exactly when one
// of these pairs is out of order:
//   t_min  > t_max_i  (i = x, y, z)      t_min_i > t_max          (i = x, y, z)
//   t_min_i > t_max_j (i != j)
// (t_min is never compared with t_max, and t_min_i <= t_max_i always holds
// for one axis because both are computed with the same sign of inv.)  With
// NaN-ignoring max/min over the three axes, lo = max_i t_min_i and
// hi = min_i t_max_i, that is
//   fail = t_min > hi  or  lo > hi  or  lo > t_max.
// The first two do not involve t_max.  The last is monotone in t_max: a box
// that fails it fails again for any smaller t_max, and whether it passes
// later needs only lo, which does not change.
struct BoxHit
{
    bool ok;  // the t_max-independent part passed
    float lo; // entry distance: the box passes iff ok and !(lo > t_max)
};

BB_D BoxHit box_part(f3 bmin, f3 bmax, f3 o, float t_min, f3 inv, bool n0, bool n1, bool n2)
{
    const float lx = ((n0 ? bmax.x : bmin.x) - o.x) * inv.x;
    const float hx = ((n0 ? bmin.x : bmax.x) - o.x) * inv.x;
    const float ly = ((n1 ? bmax.y : bmin.y) - o.y) * inv.y;
    const float hy = ((n1 ? bmin.y : bmax.y) - o.y) * inv.y;
    const float lz = ((n2 ? bmax.z : bmin.z) - o.z) * inv.z;
    const float hz = ((n2 ? bmin.z : bmax.z) - o.z) * inv.z;
    BoxHit h;
    h.lo = fmaxf(fmaxf(lx, ly), lz);
    const float hi = fminf(fminf(hx, hy), hz);
    h.ok = !((t_min > hi) | (h.lo > hi));
    return h;
}

BB_D bool box_pass(BoxHit h, float t_max) { return h.ok && !(h.lo > t_max); }

// Triangle::Intersect's test (src/geometry/triangle.cpp:55-97) with
// e1 = p1 - p0, e2 = p2 - p0 precomputed (same float32 subtraction) and the
// direction normalised once per ray (same Vector3::Normalize result).
BB_D bool mt_test(f3 p0, f3 e1, f3 e2, f3 o, f3 dn, float l, float t_min, float t_max, float& t, float& u, float& v)
{
    f3 pvec = cross(dn, e2);
    float det = dot(e1, pvec);
    if (fabsf(det) < kEps) return false;
    float inv_det = 1.0f / det;
    f3 tvec = o - p0;
    u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    f3 qvec = cross(tvec, e1);
    v = dot(dn, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qvec) * inv_det / l;
    if (t < t_min || t > t_max) return false;
    return true;
}

// Sphere::Intersect's root search (src/geometry/sphere.cpp:35-61)
BB_D void sphere_local_ray(const BbSphere& sp, f3 o, f3 d, f3& ro, f3& rd)
{
    quat q{ sp.q[0], sp.q[1], sp.q[2], sp.q[3] };
    f3 p = mk3(sp.p[0], sp.p[1], sp.p[2]);
    f3 s = mk3(sp.s[0], sp.s[1], sp.s[2]);
    ro = quat_rotate_inv(q, o - p) / s;
    rd = quat_rotate_inv(q, d / s);
}

BB_D bool sphere_test(const BbSphere& sp, f3 o, f3 d, float t_min, float t_max, float& t)
{
    f3 ro, rd;
    sphere_local_ray(sp, o, d, ro, rd);
    float a = length2(rd);
    float half_b = dot(ro, rd);
    float c = length2(ro) - sp.radius * sp.radius;
    float disc = half_b * half_b - a * c;
    if (disc < 0.0f) return false;
    float sqrt_d = sqrtf(disc);
    float root = (-half_b - sqrt_d) / a;
    if (root < t_min || t_max < root) {
        root = (-half_b + sqrt_d) / a;
        if (root < t_min || t_max < root) return false;
    }
    t = root;
    return true;
}

struct TraceCount
{
    uint32_t nodes, prims, accepts;
};

// Per-lane traversal state of BVH::RayCast (include/bulbit/bvh.h:101-155).
// `cur` is the node to expand next, whose own box test has passed: an inner
// record (>= 0), a leaf (~first slot), or kNoRef (take the next stack entry).
constexpr int32_t kNoRef = INT32_MIN;

struct Lane
{
    f3 o, d, inv, dn;
    float l, t_min, t_max;
    bool n0, n1, n2;
    int32_t sp, cur;
    int32_t hit_slot;
    float hit_t, hit_u, hit_v;
};

// Triangle::GetTexCoord (src/geometry/triangle.cpp:31-38): w uv0 + u uv1 + v uv2
BB_D f2 tri_uv(const SceneView& S, int4 vi, float u, float v)
{
    if (!S.uvs) return mk2(0.0f, 0.0f);
    const float w = 1.0f - u - v;
    const float2 a = S.uvs[vi.x], b = S.uvs[vi.y], c = S.uvs[vi.z];
    return mk2(w * a.x + u * b.x + v * c.x, w * a.y + u * b.y + v * c.y);
}

// a hit whose candidate was accepted on a restarted ray (alpha_chain): the
// slot carries this bit in the hit record
constexpr int32_t kHitRestart = 1 << 30;
constexpr int kAlphaChain = 16; // re-intersections followed before giving up (a miss)

// Primitive::Intersect's alpha test (include/bulbit/primitive.h:59-75) for
// the triangle candidate (t, u, v) found on the ray (o, d): kept when
// alpha >= 1 or when HashFloat(ray, point) does not exceed alpha (alpha <= 0:
// never kept).  A rejected candidate re-intersects the same triangle on the
// ray (point, Normalize(d)) over [Ray::epsilon, t_max - t]: for a flat
// triangle that ray meets the plane again only through the rounding of the
// point, at grazing angles, and the test repeats there.  On a keep after a
// restart, (o, d, t, u, v) describe the restarted ray and its hit and
// `restarted` is set, as the reference's Intersection holds them.
BB_D bool alpha_chain(const SceneView& S, int32_t slot, f3& o, f3& d, float t_max, float& t, float& u, float& v,
                      bool& restarted)
{
    const int4 vi = S.slot_shade[slot];
    const GpuMaterial& m = S.materials[vi.w];
    const f3 p0 = xyz(S.tris[3 * slot]), e1 = xyz(S.tris[3 * slot + 1]), e2 = xyz(S.tris[3 * slot + 2]);
    for (int k = 0; k < kAlphaChain; ++k) {
        const float alpha = texture1(S, m.tex_alpha, tri_uv(S, vi, u, v));
        if (!(alpha < 1.0f)) return true;
        const f3 point = o + d * t;
        const float p = alpha <= 0.0f ? 1.0f : hash_float_ray_point(o, d, point);
        if (!(p > alpha)) return true;
        t_max = t_max - t;
        d = normalize(d);
        o = point;
        f3 dn = d;
        const float l = normalize_m(dn);
        if (!mt_test(p0, e1, e2, o, dn, l, kRayEps, t_max, t, u, v)) return false;
        restarted = true;
    }
    return false;
}

BB_D void lane_start(Lane& L, f3 o, f3 d, float t_min, float t_max)
{
    L.o = o;
    L.d = d;
    L.t_min = t_min;
    L.t_max = t_max;
    L.inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    L.n0 = L.inv.x < 0.0f;
    L.n1 = L.inv.y < 0.0f;
    L.n2 = L.inv.z < 0.0f;
    L.dn = d;
    L.l = normalize_m(L.dn);
    L.sp = 0;
    L.cur = kNoRef;
    L.hit_slot = -1;
    L.hit_t = L.hit_u = L.hit_v = 0.0f;
}

// the root is the first node RayCast takes off its stack and box-tests
template <bool COUNT>
BB_D void lane_root(const SceneView& S, Lane& L, TraceCount& c)
{
    if (COUNT) c.nodes++;
    const BoxHit h = box_part(mk3(S.root_box[0], S.root_box[1], S.root_box[2]),
                              mk3(S.root_box[3], S.root_box[4], S.root_box[5]), L.o, L.t_min, L.inv, L.n0, L.n1,
                              L.n2);
    L.cur = box_pass(h, L.t_max) ? (COUNT ? S.root_ref : S.quad_root) : kNoRef;
}

// Traversal stack: entries [0, lds_entries) in LDS ([entry][thread] layout,
// conflict-free for a wave), deeper entries in a global overflow area
// ([entry][global thread]); cap = BVH depth + 2 bounds every push.  An entry
// is a (node reference, entry distance) pair (int2), or the reference alone
// (int) for any-hit queries without counters, whose t_max never changes
// while they run, so the pop-time `lo > t_max` check cannot fail.
template <typename E>
struct TraceStack
{
    E* lds;
    E* gbl;
    int32_t lds_entries;
    int32_t lstride; // LDS entries of one lane are this far apart (the block size)
    int32_t gstride;
    int32_t cap;
};

// stack entries of a query kind
template <bool ANY, bool COUNT>
using StackEntry = typename std::conditional<ANY && !COUNT, int, int2>::type;

// Where a lane's traversal stack lives (the SK template parameter):
//   kStackSplit    the first lds_entries entries in LDS, deeper ones in the
//                  global overflow area
//   kStackLds      the whole stack in LDS (cap <= lds_entries): no overflow branch
//   kStackScratch  the whole stack in a per-lane private array (scratch
//                  memory, cached by L1/L2), no LDS at all
enum : int
{
    kStackSplit = 0,
    kStackLds = 1,
    kStackScratch = 2
};
constexpr int kScratchStackEntries = 64; // cap of the private-array stack

template <int SK, typename E>
BB_D void stack_put(const TraceStack<E>& s, int32_t sp, E v)
{
    if (SK == kStackScratch)
        s.lds[sp] = v;
    else if (SK == kStackLds || sp < s.lds_entries)
        s.lds[sp * s.lstride] = v;
    else
        s.gbl[(sp - s.lds_entries) * s.gstride] = v;
}

template <int SK, typename E>
BB_D E stack_get(const TraceStack<E>& s, int32_t sp)
{
    if (SK == kStackScratch) return s.lds[sp];
    return (SK == kStackLds || sp < s.lds_entries) ? s.lds[sp * s.lstride] : s.gbl[(sp - s.lds_entries) * s.gstride];
}

// a far child whose t_max-independent test failed, kept on the stack only
// when counting so that it is counted when RayCast would pop and test it
// (a signalling-NaN bit pattern: no arithmetic result has it)
constexpr int32_t kFailedLo = 0x7fbfffff;

// One step of RayCast for one lane.
//
// Expanding an inner record tests both children's boxes from the parent's
// 64 bytes.  RayCast pushes the far child, then tests the near child with
// the current t_max: the near test here is that same test.  The far child is
// tested by RayCast when popped, with the t_max of that moment; here its
// t_max-independent part and entry distance are computed now, a child that
// fails already is not pushed (it would fail then too), and the pop applies
// the remaining `lo > t_max` check with the then-current t_max.  So the same
// nodes are expanded in the same order with the same outcomes, and leaves
// test their primitives in stored order (the callbacks of BVH::Intersect /
// IntersectAny, src/accel/bvh.cpp:299-361).  With COUNT, every far child is
// pushed and counted when popped, which reproduces RayCast's node-test count.
// Returns true when the lane's query is finished.
// An entry of a two-level record: its reference and box test
struct QEnt
{
    int32_t ref;
    BoxHit h;
};

BB_D QEnt qsel(bool c, QEnt a, QEnt b)
{
    QEnt r;
    r.ref = c ? a.ref : b.ref;
    r.h.ok = c ? a.h.ok : b.h.ok;
    r.h.lo = c ? a.h.lo : b.h.lo;
    return r;
}

// Expands a two-level record (SceneView::quads) for one lane.  RayCast at
// the record's node N tests N's near child now and pushes the far one; at
// an inner child it again tests the near grandchild now and pushes the far
// one.  Here the up-to-four entries (grandchildren, or a child that is a
// leaf) are ordered the same way from the ray's direction signs: near
// child's near and far entry, then far child's near and far entry.  The
// first is tested now and the others are pushed in reverse, each with its
// entry distance for the pop-time `lo > t_max` check.  That visits the same
// leaves in the same order with the same t_max as RayCast: no leaf lies
// between a child's test and its near grandchild's, so both meet the same
// t_max, and a grandchild's box lies inside its parent's, so its slab
// interval lies inside the parent's (the float slab distances are monotone
// in the plane coordinates, NaN counting as ignored) and it fails whenever
// the parent would.  Only the node-test count differs, so the counting
// kernels keep the one-level records.
template <bool ANY, int SK>
BB_D bool quad_expand(const SceneView& S, Lane& L, const TraceStack<StackEntry<ANY, false>>& stk, int* err)
{
    constexpr bool kSlim = std::is_same<StackEntry<ANY, false>, int>::value;
    const float4* r = S.quads + 8 * L.cur;
    const float4 q0 = r[0], q1 = r[1], q2 = r[2], q3 = r[3], q4 = r[4], q5 = r[5], q6 = r[6], q7 = r[7];
    // the record's 128 bytes in one round trip
    asm volatile("" ::"v"(q0.x), "v"(q0.y), "v"(q0.z), "v"(q0.w), "v"(q1.x), "v"(q1.y), "v"(q1.z), "v"(q1.w),
                 "v"(q2.x), "v"(q2.y), "v"(q2.z), "v"(q2.w), "v"(q3.x), "v"(q3.y), "v"(q3.z), "v"(q3.w), "v"(q4.x),
                 "v"(q4.y), "v"(q4.z), "v"(q4.w), "v"(q5.x), "v"(q5.y), "v"(q5.z), "v"(q5.w), "v"(q6.x), "v"(q6.y),
                 "v"(q6.z), "v"(q6.w), "v"(q7.x));
    const int32_t meta = __float_as_int(q7.x);
    const bool has1 = (meta >> 6) & 1, has3 = (meta >> 7) & 1;
    QEnt e0, e1, e2, e3;
    e0.ref = __float_as_int(q6.x);
    e1.ref = __float_as_int(q6.y);
    e2.ref = __float_as_int(q6.z);
    e3.ref = __float_as_int(q6.w);
    e0.h = box_part(mk3(q0.x, q0.y, q0.z), mk3(q0.w, q1.x, q1.y), L.o, L.t_min, L.inv, L.n0, L.n1, L.n2);
    e1.h = box_part(mk3(q1.z, q1.w, q2.x), mk3(q2.y, q2.z, q2.w), L.o, L.t_min, L.inv, L.n0, L.n1, L.n2);
    e2.h = box_part(mk3(q3.x, q3.y, q3.z), mk3(q3.w, q4.x, q4.y), L.o, L.t_min, L.inv, L.n0, L.n1, L.n2);
    e3.h = box_part(mk3(q4.z, q4.w, q5.x), mk3(q5.y, q5.z, q5.w), L.o, L.t_min, L.inv, L.n0, L.n1, L.n2);
    e1.h.ok = e1.h.ok && has1;
    e3.h.ok = e3.h.ok && has3;
    const int32_t aN = meta & 3, aA = (meta >> 2) & 3, aB = (meta >> 4) & 3;
    const bool nN = aN == 0 ? L.n0 : (aN == 1 ? L.n1 : L.n2);
    const bool nA = aA == 0 ? L.n0 : (aA == 1 ? L.n1 : L.n2);
    const bool nB = aB == 0 ? L.n0 : (aB == 1 ? L.n1 : L.n2);
    // each child's entries in near-far order (child2 is near when d[axis] < 0)
    const bool sA = has1 && nA, sB = has3 && nB;
    const QEnt a_near = qsel(sA, e1, e0), a_far = qsel(sA, e0, e1);
    const QEnt b_near = qsel(sB, e3, e2), b_far = qsel(sB, e2, e3);
    const QEnt x0 = qsel(nN, b_near, a_near), x1 = qsel(nN, b_far, a_far);
    const QEnt x2 = qsel(nN, a_near, b_near), x3 = qsel(nN, a_far, b_far);
    if (L.sp + 3 > stk.cap) {
        atomicOr(err, 1);
        return true;
    }
    // pushed only when they pass with the current t_max (t_max only shrinks)
    if (box_pass(x3.h, L.t_max)) {
        if constexpr (kSlim)
            stack_put<SK>(stk, L.sp, x3.ref);
        else
            stack_put<SK>(stk, L.sp, make_int2(x3.ref, __float_as_int(x3.h.lo)));
        ++L.sp;
    }
    if (box_pass(x2.h, L.t_max)) {
        if constexpr (kSlim)
            stack_put<SK>(stk, L.sp, x2.ref);
        else
            stack_put<SK>(stk, L.sp, make_int2(x2.ref, __float_as_int(x2.h.lo)));
        ++L.sp;
    }
    if (box_pass(x1.h, L.t_max)) {
        if constexpr (kSlim)
            stack_put<SK>(stk, L.sp, x1.ref);
        else
            stack_put<SK>(stk, L.sp, make_int2(x1.ref, __float_as_int(x1.h.lo)));
        ++L.sp;
    }
    if (box_pass(x0.h, L.t_max)) {
        L.cur = x0.ref;
        return false;
    }
    L.cur = kNoRef;
    return false;
}

template <bool ANY, bool COUNT, int SK, bool ALPHA>
BB_D bool trace_step(const SceneView& S, Lane& L, const TraceStack<StackEntry<ANY, COUNT>>& stk, TraceCount& c,
                     int* err)
{
    constexpr bool kSlim = std::is_same<StackEntry<ANY, COUNT>, int>::value;
    if (!COUNT && L.cur >= 0) {
        // two-level records (no node counters kept)
        if constexpr (!COUNT) {
            if (quad_expand<ANY, SK>(S, L, stk, err)) return true;
            if (L.cur != kNoRef) return false;
        }
    } else if (L.cur >= 0) {
        const float4* r = S.inner + 4 * L.cur;
        const float4 a = r[0], b = r[1], e = r[2], m = r[3];
        // the record's 64 bytes in one round trip
        asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b.x), "v"(b.y), "v"(b.z), "v"(b.w), "v"(e.x),
                     "v"(e.y), "v"(e.z), "v"(e.w), "v"(m.x), "v"(m.y), "v"(m.z));
        const BoxHit h1 = box_part(mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), L.o, L.t_min, L.inv, L.n0, L.n1, L.n2);
        const BoxHit h2 = box_part(mk3(b.z, b.w, e.x), mk3(e.y, e.z, e.w), L.o, L.t_min, L.inv, L.n0, L.n1, L.n2);
        const int32_t axis = __float_as_int(m.z);
        const bool neg = axis == 0 ? L.n0 : (axis == 1 ? L.n1 : L.n2);
        const int32_t near_ref = __float_as_int(neg ? m.y : m.x);
        const int32_t far_ref = __float_as_int(neg ? m.x : m.y);
        const BoxHit hn = neg ? h2 : h1;
        const BoxHit hf = neg ? h1 : h2;
        const bool push = COUNT || box_pass(hf, L.t_max);
        if (push) {
            if (L.sp >= stk.cap) {
                atomicOr(err, 1);
                return true;
            }
            if constexpr (kSlim)
                stack_put<SK>(stk, L.sp, far_ref);
            else
                stack_put<SK>(stk, L.sp, make_int2(far_ref, hf.ok ? __float_as_int(hf.lo) : kFailedLo));
            ++L.sp;
        }
        if (COUNT) c.nodes++;
        if (box_pass(hn, L.t_max)) {
            L.cur = near_ref;
            return false;
        }
    } else if (L.cur != kNoRef) {
        for (int32_t slot = ~L.cur;; ++slot) {
            const float4 t0 = S.tris[3 * slot];
            const float4 t1 = S.tris[3 * slot + 1];
            const float4 t2 = S.tris[3 * slot + 2];
            // keep the three 16-byte loads together ahead of the first use
            asm volatile("" ::"v"(t0.x), "v"(t0.y), "v"(t0.z), "v"(t1.x), "v"(t1.y), "v"(t1.z), "v"(t2.x), "v"(t2.y),
                         "v"(t2.z), "v"(t2.w));
            if (COUNT) c.prims++;
            const int32_t kind = __float_as_int(t2.w);
            float t = 0.0f, u = 0.0f, v = 0.0f;
            // the triangle test runs on every slot so that the three loads
            // above issue together; a sphere slot (zero edges) fails it and
            // takes the sphere test instead
            bool h = mt_test(xyz(t0), xyz(t1), xyz(t2), L.o, L.dn, L.l, L.t_min, L.t_max, t, u, v);
            if ((kind & KIND_MASK) == KIND_SPHERE) {
                u = v = 0.0f;
                h = sphere_test(S.spheres[__float_as_int(t1.w)], L.o, L.d, L.t_min, L.t_max, t);
            }
            bool restarted = false;
            if constexpr (ALPHA) { // scenes with an alpha texture only
                if (h && (kind & kSlotAlpha)) {
                    f3 ao = L.o, ad = L.d;
                    h = alpha_chain(S, slot, ao, ad, L.t_max, t, u, v, restarted);
                }
            }
            if (h) {
                if (COUNT) c.accepts++;
                L.hit_slot = slot | (restarted ? kHitRestart : 0);
                if (ANY) return true;
                L.hit_t = t;
                L.hit_u = u;
                L.hit_v = v;
                if (t <= L.t_min) return true;
                L.t_max = t;
            }
            if (kind & kSlotLast) break;
        }
    }
    // take the next stack entry; its deferred `lo > t_max` check decides
    // whether it is expanded
    if (L.sp == 0) return true;
    --L.sp;
    if constexpr (kSlim) {
        L.cur = stack_get<SK>(stk, L.sp); // pushed only if it passed with this same t_max
    } else {
        const int2 e = stack_get<SK>(stk, L.sp);
        if (COUNT) c.nodes++;
        const bool pass = (!COUNT || e.y != kFailedLo) && !(__int_as_float(e.y) > L.t_max);
        L.cur = pass ? e.x : kNoRef;
    }
    return false;
}

// Persistent traversal: a wave takes consecutive queries from the shared
// counter `fetch` and hands a new query to its idle lanes once at least
// `refill_min` of its 64 lanes are idle (64: whole-wave batches, one atomic
// per 64 queries).  Between refill checks every lane takes STEPS traversal
// steps.  Src provides load(index, Lane&) and finish(index, const Lane&).
template <bool ANY, bool COUNT, int SK, bool ALPHA, int STEPS, typename Src>
BB_D void persistent_trace(const SceneView& S, const Src& src, int n, int* fetch,
                           const TraceStack<StackEntry<ANY, COUNT>>& stk, TraceCount& c, int* err, int refill_min)
{
    const int lane = __lane_id();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    Lane L;
    bool active = false, more = true;
    int id = 0;
    // the wave's reservoir of claimed queries [rb, re): claimed kReservoir
    // at a time, so the shared counter sees one atomic per kReservoir
    // queries instead of one per refill (same-address atomics limit the rate)
    int rb = 0, re = 0;
    for (;;) {
        unsigned long long idle = __ballot(active ? 0 : 1);
        if ((more || rb < re) && (__popcll(idle) >= refill_min || idle == ~0ull)) {
            if (rb >= re) {
                int base = 0;
                if (lane == 0) base = atomicAdd(fetch, kReservoir);
                base = __shfl(base, 0);
                rb = base;
                re = base + kReservoir < n ? base + kReservoir : n;
                if (base + kReservoir >= n) more = false;
            }
            const int rank = __popcll(idle & below);
            const int mine = rb + rank;
            if (!active && mine < re) {
                id = mine;
                src.load(id, L);
                lane_root<COUNT>(S, L, c);
                active = true;
            }
            const int avail = re - rb, need = __popcll(idle);
            rb += need < avail ? need : avail;
            idle = __ballot(active ? 0 : 1);
        }
        if (idle == ~0ull && !more && rb >= re) break;
        if (idle == ~0ull) continue;
#pragma unroll
        for (int k = 0; k < STEPS; ++k) {
            if (active && trace_step<ANY, COUNT, SK, ALPHA>(S, L, stk, c, err)) {
                src.finish(id, L);
                active = false;
            }
        }
    }
}

// wave-level compaction: returns this lane's slot (or -1) in the queue whose
// length is *counter; every lane of the wave must call it
BB_D int wave_alloc(bool pred, int* counter)
{
    const unsigned long long mask = __ballot(pred ? 1 : 0);
    if (mask == 0ull) return -1;
    const int lane = __lane_id();
    const int leader = __builtin_ctzll(mask);
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long below = (lane == 0) ? 0ull : (mask & (~0ull >> (64 - lane)));
    return pred ? base + __popcll(below) : -1;
}

// Wave-level compaction among the lanes that reach the call together
// (divergent code allowed): one atomic per wave; returns this lane's slot
BB_D int wave_alloc_active(int* counter)
{
    const unsigned long long mask = __ballot(1);
    const int lane = __lane_id();
    const int leader = __builtin_ctzll(mask);
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return base + __popcll(mask & below);
}

// Block-level compaction into two queues at once: one atomic per queue per
// block instead of per wave (same-address atomics were the shading kernel's
// limit).  Every thread of the block must call it; s holds
// 2 * (waves per block) + 2 ints of LDS.  Slots: a = this thread's place in
// the queue counted by ca (when pa), b likewise.
BB_D void block_alloc2(bool pa, int* ca, int& a, bool pb, int* cb, int& b, int* s)
{
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = __lane_id();
    const unsigned long long ma = __ballot(pa ? 1 : 0), mb = __ballot(pb ? 1 : 0);
    if (lane == 0) {
        s[wave] = __popcll(ma);
        s[nw + wave] = __popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tb = 0;
        for (int w = 0; w < nw; ++w) {
            const int x = s[w], y = s[nw + w];
            s[w] = ta;
            s[nw + w] = tb;
            ta += x;
            tb += y;
        }
        s[2 * nw] = ta ? atomicAdd(ca, ta) : 0;
        s[2 * nw + 1] = tb ? atomicAdd(cb, tb) : 0;
    }
    __syncthreads();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    a = pa ? s[2 * nw] + s[wave] + __popcll(ma & below) : -1;
    b = pb ? s[2 * nw + 1] + s[nw + wave] + __popcll(mb & below) : -1;
    __syncthreads(); // s is reused by the next call
}

BB_D void wave_count_add(TraceCount c, unsigned long long* out)
{
    unsigned long long n = c.nodes, p = c.prims, a = c.accepts;
    for (int off = 32; off > 0; off >>= 1) {
        n += __shfl_xor(n, off);
        p += __shfl_xor(p, off);
        a += __shfl_xor(a, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(&out[0], n);
        atomicAdd(&out[1], p);
        atomicAdd(&out[2], a);
    }
}

// ------------------------------------------------------- batch traversal

struct BatchSrc
{
    const float4* rays;
    int4* hits;
    const float4* tris;
    int any_hit;
    BB_D void load(int i, Lane& L) const
    {
        const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
        lane_start(L, xyz(r0), xyz(r1), r0.w, r1.w);
    }
    BB_D void finish(int i, const Lane& L) const
    {
        int4 out;
        if (any_hit)
            out = make_int4(L.hit_slot >= 0 ? 1 : -1, 0, 0, 0);
        else if (L.hit_slot >= 0)
            out = make_int4(__float_as_int(tris[3 * (L.hit_slot & ~kHitRestart)].w), __float_as_int(L.hit_t),
                            __float_as_int(L.hit_u),
                            __float_as_int(L.hit_v));
        else
            out = make_int4(-1, 0, 0, 0);
        hits[i] = out;
    }
};

// most entries a traversal stack holds: one-level records (with node
// counters) or two-level records
BB_HD int trace_stack_depth(const SceneView& V, bool count) { return count ? V.stack_depth : V.quad_stack_depth; }

// the calling thread's stack in the block's dynamic LDS and in the overflow
// area (both hold entries of type E)
template <typename E>
BB_D TraceStack<E> make_stack(void* lds_stack, int2* gstack, int lds_entries, int cap)
{
    TraceStack<E> t;
    t.lds = reinterpret_cast<E*>(lds_stack) + threadIdx.x;
    t.lstride = blockDim.x;
    t.gstride = gridDim.x * blockDim.x;
    t.gbl = reinterpret_cast<E*>(gstack) + blockIdx.x * blockDim.x + threadIdx.x;
    t.lds_entries = lds_entries;
    t.cap = cap;
    return t;
}

// the calling lane's stack in its private array (kStackScratch)
template <typename E>
BB_D TraceStack<E> make_scratch_stack(E* priv, int cap)
{
    TraceStack<E> t;
    t.lds = priv;
    t.lstride = 1;
    t.gstride = 0;
    t.gbl = nullptr;
    t.lds_entries = kScratchStackEntries;
    t.cap = cap < kScratchStackEntries ? cap : kScratchStackEntries;
    return t;
}

#ifndef BB_TRACE_STEPS
#define BB_TRACE_STEPS 2
#endif
constexpr int kTraceSteps = BB_TRACE_STEPS;
#ifndef BB_SHADOW_STEPS
#define BB_SHADOW_STEPS 2
#endif
constexpr int kShadowSteps = BB_SHADOW_STEPS;

template <bool ANY, bool COUNT, int SK, bool ALPHA>
__global__ __launch_bounds__(kTraceBlock) void k_trace_batch(SceneView S, BatchSrc src, int n, int* fetch,
                                                               int2* gstack, int lds_entries, int refill_min,
                                                               unsigned long long* counters, int* err)
{
    extern __shared__ int2 lds_stack[];
    const auto stk = make_stack<StackEntry<ANY, COUNT>>(lds_stack, gstack, lds_entries, trace_stack_depth(S, COUNT));
    TraceCount c{ 0, 0, 0 };
    persistent_trace<ANY, COUNT, SK, ALPHA, kTraceSteps>(S, src, n, fetch, stk, c, err, refill_min);
    if (COUNT) wave_count_add(c, counters);
}

// -------------------------------------------------------- path tracer state

constexpr int kFilterN = 32; // GaussianFilter table size (gaussian_filter.cpp:10)

struct GpuCamera
{
    f3 origin, lower_left, horizontal, vertical, u, v;
    float lens_radius, filter_extent;
    int32_t width, height;
    int32_t filter_type;
    const float* filter_cdf; // GAUSSIAN: marginal CDF, then 32 conditional CDFs (33 floats each)
};

// Distribution1D::SampleContinuous (sampling.h:381-398) over a CDF of n + 1
// entries, with FindInterval (math.h:117-135)
BB_D float cdf_sample(const float* cdf, int32_t n, float u, int32_t& off)
{
    int32_t first = 0, len = n + 1;
    while (len > 0) {
        const int32_t half = len >> 1, middle = first + half;
        if (cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    off = min(max(first - 1, 0), n - 1);
    const float du = (u - cdf[off]) / (cdf[off + 1] - cdf[off]);
    return (float(off) + du) / float(n);
}

// Filter::Sample: box_filter.cpp:17-21, tent_filter.cpp:22-30,
// GaussianFilter (gaussian_filter.cpp:40-43 over Distribution2D::
// SampleContinuous, sampling.h:447-461)
BB_D f2 filter_sample(const GpuCamera& cam, f2 u)
{
    const float e = cam.filter_extent;
    if (cam.filter_type == BB_FILTER_TENT) {
        const float h = e / 2.0f;
        const float x = u.x < 0.5f ? h * (sqrtf(2.0f * u.x) - 1.0f) : h * (1.0f - sqrtf(1.0f - 2.0f * (u.x - 0.5f)));
        const float y = u.y < 0.5f ? h * (sqrtf(2.0f * u.y) - 1.0f) : h * (1.0f - sqrtf(1.0f - 2.0f * (u.y - 0.5f)));
        return mk2(x, y);
    }
    if (cam.filter_type == BB_FILTER_GAUSSIAN) {
        int32_t v = 0, unused = 0;
        const float d1 = cdf_sample(cam.filter_cdf, kFilterN, u.y, v);
        const float d0 = cdf_sample(cam.filter_cdf + (kFilterN + 1) * (1 + v), kFilterN, u.x, unused);
        const float h = e / 2.0f;
        return mk2(d0 * e - h, d1 * e - h);
    }
    return mk2((u.x * 2.0f - 1.0f) * e / 2.0f, (u.y * 2.0f - 1.0f) * e / 2.0f);
}

struct PathBufs
{
    float4* L;     // L.rgb, eta_scale
    float4* beta;  // beta.rgb, prev_bsdf_pdf
    uint64_t* rng; // PCG32 state
    int32_t* flags;  // bounce | specular << 16
    int32_t* queue[2];
    float4* ray[2];  // 2 float4 per queue slot: (o, -), (d, -)
    int4* hit;       // (slot, t, u, v) per queue slot
    float4* sray;    // 2 float4 per shadow slot: (o, t_max), (d, path id)
    float4* scontrib;
    int* counts;     // [depth][4]: active paths, shadow rays, closest fetch, shadow fetch
    int2* gstack;    // traversal stack overflow, [entry][persistent thread]
    int* err;
    unsigned long long* counters; // [6]: closest nodes/prims/accepts, shadow nodes/prims/accepts
    int* sample_begin;            // the current call's sample_begin
};

struct RenderConsts
{
    int32_t n_paths, npix, width, seed, max_bounces, rr_min_bounces;
    int32_t sample0;              // the pass's first sample relative to the call's sample_begin
    const int32_t* sample_begin;  // device copy of the call's sample_begin (set per call, read by k_raygen)
    int32_t train_stride; // cached rendering: every train_stride-th path is a training path (0: none)
    int32_t light_alias;  // 1: alias-table light selection (BB_LIGHTS_ALIAS)
};

// Path flags: bounce (bits 0-15), the last BSDF sample was specular (16),
// training path (18), and for a training path the vertex at which it stops
// (bits 20-27, 255 = not yet known)
constexpr int32_t kFlagSpecular = 1 << 16;
constexpr int32_t kFlagTraining = 1 << 18;
constexpr int32_t kStopShift = 20;
constexpr int32_t kStopUnknown = 255;

// Buffers of the cached path tracer (radiance cache queries and training
// vertices); raw feature rows are the cache's 16-float inputs
struct CacheBufs
{
    float* qraw;      // [query][16]
    float4* qbeta;    // [query] throughput at the query vertex, path id
    float* qpred;     // [query][3]
    float* traw;      // [training path][vertex][16]
    float4* tsnap;    // [training path][vertex][2]: L after the vertex's emission, throughput
    int32_t* tcount;  // [training path] recorded vertices
    float* rows;      // [row][16] compacted training rows
    float* targets;   // [row][3]
    int32_t* counters; // [0] queries, [1] training rows
    int32_t* diverged; // the network's divergence flag (sticky)
    int32_t* state;    // [0] diverged when the pass began, [1] a prediction of this pass was not finite,
                       // [2] 1 + the pass dropped for it (0: none)
    unsigned long long* tally; // [0] queries, [1] rows trained (summed over the call's passes)
    f3 wmin, winv;    // position normalisation to [0, 1]^3
    int32_t query_bounce, query_first, query_nee, extra, max_vertices, max_rows, max_queries;
};

// UniDirectionalRayIntegrator::Render per (pixel, sample) start
// (integrator.cpp:187-190): seed, then SampleRay(Next2D(), Next2D()) with the
// lens sample u1 drawn first (argument evaluation of the reference's compiler),
// PerspectiveCamera::SampleRay (perspective_camera.cpp:40-52) with the
// camera's Filter::Sample.
__global__ __launch_bounds__(kShadeBlock) void k_raygen(GpuCamera cam, RenderConsts rc, PathBufs P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        P.counts[0] = rc.n_paths;
    }
    if (i >= rc.n_paths) return;
    const int pix = i % rc.npix;
    const int sample = *rc.sample_begin + rc.sample0 + i / rc.npix;
    const int px = pix % rc.width, py = pix / rc.width;
    uint64_t st = pcg_seed(pixel_sample_hash(px, py, rc.seed, sample));
    f2 u1;
    u1.x = pcg_float(st);
    u1.y = pcg_float(st);
    f2 u0;
    u0.x = pcg_float(st);
    u0.y = pcg_float(st);
    f2 off = filter_sample(cam, u0);
    off = mk2(off.x + 0.5f, off.y + 0.5f);
    f3 pixel_center = (cam.lower_left + (cam.horizontal * (float(px) + off.x)) / float(cam.width)) +
                      (cam.vertical * (float(py) + off.y)) / float(cam.height);
    f2 ap = sample_concentric(u1);
    ap = mk2(ap.x * cam.lens_radius, ap.y * cam.lens_radius);
    f3 offset = cam.u * ap.x + cam.v * ap.y;
    f3 o = cam.origin + offset;
    f3 d = normalize(pixel_center - o);
    P.ray[0][2 * i] = make_float4(o.x, o.y, o.z, 0.0f);
    P.ray[0][2 * i + 1] = make_float4(d.x, d.y, d.z, 0.0f);
    P.queue[0][i] = i;
    P.L[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    P.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    P.rng[i] = st;
    P.flags[i] = (rc.train_stride > 0 && i % rc.train_stride == 0) ? (kFlagTraining | (kStopUnknown << kStopShift)) : 0;
}

struct ClosestSrc
{
    const float4* rays;
    int4* hits;
    BB_D void load(int i, Lane& L) const
    {
        const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
        lane_start(L, xyz(r0), xyz(r1), kRayEps, INFINITY);
    }
    BB_D void finish(int i, const Lane& L) const
    {
        hits[i] = make_int4(L.hit_slot, __float_as_int(L.hit_t), __float_as_int(L.hit_u), __float_as_int(L.hit_v));
    }
};

// Shadow rays of SampleDirectLight (path_integrator.cpp:178-193): an
// unoccluded ray adds its precomputed contribution to the path's L
// (L += SampleDirectLight(...), path_integrator.cpp:109-112).
struct ShadowSrc
{
    const float4* srays;
    const float4* contrib;
    float4* L;
    BB_D void load(int i, Lane& Ln) const
    {
        const float4 r0 = srays[2 * i], r1 = srays[2 * i + 1];
        lane_start(Ln, xyz(r0), xyz(r1), kRayEps, r0.w);
    }
    BB_D void finish(int i, const Lane& Ln) const
    {
        if (Ln.hit_slot >= 0) return;
        const int pid = __float_as_int(srays[2 * i + 1].w);
        const float4 cv = contrib[i];
        float4 l = L[pid];
        l.x = l.x + cv.x;
        l.y = l.y + cv.y;
        l.z = l.z + cv.z;
        L[pid] = l;
    }
};

// the closest-hit and shadow kernels of the bounce loop, for blocks of
// BLOCK threads and the stack placement SK
template <bool COUNT, int SK, bool ALPHA, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_closest(SceneView S, const int* __restrict__ n_ptr, int* fetch,
                                                         ClosestSrc src, int2* gstack, int lds_entries, int refill_min,
                                                         unsigned long long* counters, int* err)
{
    using E = StackEntry<false, COUNT>;
    extern __shared__ int2 lds_stack[];
    E priv[SK == kStackScratch ? kScratchStackEntries : 1];
    const auto stk = SK == kStackScratch
                         ? make_scratch_stack<E>(priv, trace_stack_depth(S, COUNT))
                         : make_stack<E>(lds_stack, gstack, lds_entries, trace_stack_depth(S, COUNT));
    TraceCount c{ 0, 0, 0 };
    persistent_trace<false, COUNT, SK, ALPHA, kTraceSteps>(S, src, *n_ptr, fetch, stk, c, err, refill_min);
    if (COUNT) wave_count_add(c, counters);
}

template <bool COUNT, int SK, bool ALPHA, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_shadow(SceneView S, const int* __restrict__ n_ptr, int* fetch,
                                                        ShadowSrc src, int2* gstack, int lds_entries, int refill_min,
                                                        unsigned long long* counters, int* err)
{
    using E = StackEntry<true, COUNT>;
    extern __shared__ int2 lds_stack[];
    E priv[SK == kStackScratch ? kScratchStackEntries : 1];
    const auto stk = SK == kStackScratch
                         ? make_scratch_stack<E>(priv, trace_stack_depth(S, COUNT))
                         : make_stack<E>(lds_stack, gstack, lds_entries, trace_stack_depth(S, COUNT));
    TraceCount c{ 0, 0, 0 };
    persistent_trace<true, COUNT, SK, ALPHA, kShadowSteps>(S, src, *n_ptr, fetch, stk, c, err, refill_min);
    if (COUNT) wave_count_add(c, counters + 3);
}

// Surface data of the closest hit: the shading part of Triangle::Intersect
// (triangle.cpp:99-107) / Sphere::Intersect (sphere.cpp:63-82) and
// Shape::SetFaceNormal (shape.h:32-43).
struct SurfaceHit
{
    f3 point, normal, ns, ts;
    f3 nl; // sphere: the object-space normal
    bool front, sphere;
};

BB_D SurfaceHit surface_at(const SceneView& S, int32_t slot, float t, float u, float v, f3 o, f3 d)
{
    SurfaceHit h;
    const float4 t0 = S.tris[3 * slot], t1 = S.tris[3 * slot + 1], t2 = S.tris[3 * slot + 2];
    f3 normal, ns, ts;
    if ((__float_as_int(t2.w) & KIND_MASK) == KIND_SPHERE) {
        const BbSphere& sp = S.spheres[__float_as_int(t1.w)];
        quat q{ sp.q[0], sp.q[1], sp.q[2], sp.q[3] };
        f3 ro, rd;
        sphere_local_ray(sp, o, d, ro, rd);
        f3 pl = ro + rd * t;
        f3 nl = normalize(pl / sp.radius);
        h.nl = nl;
        h.sphere = true;
        h.point = quat_rotate(q, mk3(sp.s[0], sp.s[1], sp.s[2]) * pl) + mk3(sp.p[0], sp.p[1], sp.p[2]);
        f3 n = quat_rotate(q, nl);
        f3 tt = (n.y < 1.0f - 1e-4f) ? normalize(cross(mk3(0.0f, 1.0f, 0.0f), n)) : normalize(cross(n, mk3(1.0f, 0.0f, 0.0f)));
        normal = n;
        ns = n;
        ts = tt;
    } else {
        h.sphere = false;
        h.nl = mk3(0.0f);
        const int4 vi = S.slot_shade[slot];
        const f3 e1 = xyz(t1), e2 = xyz(t2);
        const float w = 1.0f - u - v;
        h.point = o + d * t;
        normal = normalize(cross(e1, e2));
        const f3 n0 = xyz(S.normals[vi.x]), n1 = xyz(S.normals[vi.y]), n2 = xyz(S.normals[vi.z]);
        const f3 g0 = xyz(S.tangents[vi.x]), g1 = xyz(S.tangents[vi.y]), g2 = xyz(S.tangents[vi.z]);
        ns = normalize(n0 * w + n1 * u + n2 * v);
        ts = normalize(g0 * w + g1 * u + g2 * v);
    }
    h.front = dot(d, normal) < 0.0f;
    const float sign = h.front ? 1.0f : -1.0f;
    h.normal = normal * sign;
    h.ns = ns * sign;
    h.ts = ts * sign;
    return h;
}

// the hit's texture coordinates: Triangle::GetTexCoord, or Sphere's
// ComputeTexCoord of the object-space normal (src/geometry/sphere.cpp:9-16)
BB_D f2 hit_uv(const SceneView& S, int32_t slot, const SurfaceHit& h, float u, float v)
{
    if (h.sphere) {
        const float theta = (float)::acos((double)clampf_(h.nl.y, -1.0f, 1.0f));
        const float r = (float)::atan2((double)h.nl.z, (double)h.nl.x);
        const float phi = r < 0.0f ? r + kTwoPi : r;
        return mk2(phi * kInvTwoPi, 1.0f - theta * kInvPi);
    }
    return tri_uv(S, S.slot_shade[slot], u, v);
}

// The hit's material with its textures evaluated at the hit's texture
// coordinates, and the normal map applied to the shading frame
BB_D void textured_hit(const SceneView& S, int32_t slot, float u, float v, f3 wo, GpuMaterial& mat,
                                          SurfaceHit& sh)
{
    const f2 uv = hit_uv(S, slot, sh, u, v);
    if (mat.textured) mat = material_at(mat, S, uv);
    if (mat.tex_normal >= 0) normal_map(S, mat.tex_normal, uv, wo, sh.normal, sh.ns, sh.ts);
}

// PathIntegrator::Li, one bounce (path_integrator.cpp:38-147) and the
// unoccluded part of SampleDirectLight (:152-194).
// the cache's input row at a surface vertex: position in [0, 1]^3, direction
// towards the previous vertex, shading normal, roughness, albedo
BB_D void cache_features(const CacheBufs& C, const GpuMaterial& m, f3 p, f3 wo, f3 ns, float* r)
{
    const f3 q = (p - C.wmin) * C.winv;
    r[0] = q.x;
    r[1] = q.y;
    r[2] = q.z;
    r[3] = wo.x;
    r[4] = wo.y;
    r[5] = wo.z;
    r[6] = ns.x;
    r[7] = ns.y;
    r[8] = ns.z;
    r[9] = m.kind == MK_LAMBERT ? 1.0f : sqrtf(fmaxf(m.ax, m.ay));
    const spec a = m.kind == MK_METALLIC_ROUGHNESS ? m.color : m.r;
    r[10] = a.r;
    r[11] = a.g;
    r[12] = a.b;
    r[13] = r[14] = r[15] = 0.0f;
}

// KM: the BxDF kinds compiled in (the scene's materials' kinds, or all)
template <bool CACHE, bool TEX, int KM>
__global__ __launch_bounds__(kShadeBlock) void k_shade(SceneView S, RenderConsts rc, PathBufs P, int depth, CacheBufs C)
{
    const int in = depth & 1, out = in ^ 1;
    const int n = P.counts[4 * depth];
    int* next_count = &P.counts[4 * (depth + 1)];
    int* shadow_count = &P.counts[4 * depth + 1];
    // the light CDF in LDS for the binary search of every NEE sample (the
    // same comparisons on the same values; large light sets stay in global memory)
    __shared__ float s_cdf[kLdsCdfEntries];
    const bool cdf_lds = S.n_lights + 1 <= kLdsCdfEntries;
    if (cdf_lds) {
        for (int i = threadIdx.x; i < S.n_lights + 1; i += blockDim.x) s_cdf[i] = S.light_cdf[i];
        __syncthreads();
    }
    const float* cdf = cdf_lds ? s_cdf : S.light_cdf;
    __shared__ int s_alloc[2 * (kShadeBlock / 64) + 2];
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int k = base + threadIdx.x;
        bool cont = false, shadow = false;
        int pid = 0;
        f3 new_o = mk3(0.0f), new_d = mk3(0.0f), sh_o = mk3(0.0f), sh_d = mk3(0.0f);
        float sh_tmax = 0.0f;
        spec sh_c = mks(0.0f);
        if (k < n) {
            pid = P.queue[in][k];
            const float4 r0 = P.ray[in][2 * k], r1 = P.ray[in][2 * k + 1];
            const f3 o = xyz(r0), d = xyz(r1);
            const int4 hh = P.hit[k];
            float4 Lv = P.L[pid];
            float4 Bv = P.beta[pid];
            uint64_t st = P.rng[pid];
            int32_t fl = P.flags[pid];
            int32_t bounce = fl & 0xffff;
            bool specular = (fl & kFlagSpecular) != 0;
            spec L = mks(Lv.x, Lv.y, Lv.z);
            float eta_scale = Lv.w;
            spec beta = mks(Bv.x, Bv.y, Bv.z);
            float prev_pdf = Bv.w;
            if (CACHE && C.query_nee && (fl & kFlagTraining)) {
                // the radiance gathered up to and including the previous
                // vertex's direct light: its indirect part starts here
                const int tp = pid / rc.train_stride;
                const int32_t nv = C.tcount[tp];
                if (nv > 0) {
                    const size_t rec = size_t(tp) * C.max_vertices + nv - 1;
                    if (C.tsnap[2 * rec].w == 0.0f) C.tsnap[2 * rec] = make_float4(L.r, L.g, L.b, 1.0f);
                }
            }
            // a restarted alpha hit lies on the restarted ray: replay the
            // chain (deterministic) from the original ray to recover it
            const int32_t slot = hh.x < 0 ? hh.x : (hh.x & ~kHitRestart);
            f3 so = o, sd = d;
            if (hh.x >= 0 && (hh.x & kHitRestart)) {
                const float4 q0 = S.tris[3 * slot], q1 = S.tris[3 * slot + 1], q2 = S.tris[3 * slot + 2];
                float rt = 0.0f, ru = 0.0f, rv = 0.0f;
                bool rs = false;
                f3 dn = d;
                const float l = normalize_m(dn);
                if (mt_test(xyz(q0), xyz(q1), xyz(q2), o, dn, l, kRayEps, INFINITY, rt, ru, rv))
                    alpha_chain(S, slot, so, sd, INFINITY, rt, ru, rv, rs);
            }
            if (slot < 0) {
                // escaped: infinite lights (path_integrator.cpp:41-62)
                for (int32_t k2 = 0; k2 < S.n_inf; ++k2) {
                    const GpuLight& Lt = S.lights[S.inf_lights[k2]];
                    if (bounce == 0 || specular) {
                        L = L + beta * Lt.L;
                    } else {
                        float light_pdf = kInvFourPi * Lt.pmf;
                        float mis = power_heuristic(prev_pdf, light_pdf);
                        L = L + (beta * mis) * Lt.L;
                    }
                }
            } else {
                const float t = __int_as_float(hh.y), u = __int_as_float(hh.z), v = __int_as_float(hh.w);
                SurfaceHit sh = surface_at(S, slot, t, u, v, so, sd);
                const f3 wo = normalize(-d);
                // the hit's material with its textures evaluated at the hit,
                // and the normal map applied to the shading frame (the
                // per-hit part of Intersection::GetBSDF, intersection.cpp:32-52)
                const int32_t mat_index = S.slot_shade[slot].w;
                GpuMaterial mat;
                if (mat_index >= 0) {
                    mat = S.materials[mat_index];
                    if constexpr (TEX) { // scenes with textured or normal-mapped materials only
                        if (mat.textured || mat.tex_normal >= 0) textured_hit(S, slot, u, v, wo, mat, sh);
                    }
                }
                const int32_t al = S.slot_light[slot];
                if (al >= 0) {
                    // area light emission (path_integrator.cpp:67-84)
                    const GpuLight& Lt = S.lights[al];
                    spec Le = (sh.front || Lt.two_sided) ? Lt.L : mks(0.0f);
                    if (!is_black(Le)) {
                        if (bounce == 0 || specular) {
                            L = L + beta * Le;
                        } else {
                            // Triangle::PDF(isect, ray) (triangle.cpp:247-253)
                            float dist2 = t * t * length2(d);
                            float cosine = fabsf(dot(d, sh.normal) / length(d));
                            float area = 0.5f * length(cross(xyz(S.tris[3 * slot + 1]), xyz(S.tris[3 * slot + 2])));
                            float light_pdf = (dist2 / (cosine * area)) * Lt.pmf;
                            float mis = power_heuristic(prev_pdf, light_pdf);
                            L = L + (beta * mis) * Le;
                        }
                    }
                }
                bool stop_here = false;
                bool nee_only = false; // a query vertex that takes its direct light before the cache
                if (CACHE) {
                    // the radiance cache replaces the rest of a rendering path at
                    // its query vertex; a training path continues `extra` vertices
                    // further and records every vertex it continues from
                    const int32_t mi = S.slot_shade[slot].w;
                    const bool training = (fl & kFlagTraining) != 0;
                    // query vertex: the first rough (non-specular) vertex, or
                    // vertex query_bounce at the latest
                    const bool query_here =
                        mi >= 0 && C.state[0] == 0 &&
                        (depth >= C.query_bounce ||
                         (depth >= C.query_first && (mat.flags_front & (BX_DIFFUSE | BX_GLOSSY))));
                    int32_t stop = (fl >> kStopShift) & 0xff;
                    if (training) {
                        if (query_here && stop == kStopUnknown) stop = min(depth + C.extra, 254);
                        fl = (fl & ~(0xff << kStopShift)) | (stop << kStopShift);
                        if (depth >= stop) {
                            stop_here = true;
                        } else if (bounce < rc.max_bounces && mi >= 0) {
                            const int tp = pid / rc.train_stride;
                            const int32_t nv = C.tcount[tp];
                            if (nv < C.max_vertices) {
                                const size_t rec = size_t(tp) * C.max_vertices + nv;
                                cache_features(C, mat, sh.point, wo, sh.ns, C.traw + 16 * rec);
                                // with direct light at the query vertex the cache learns the
                                // indirect part: the snapshot is taken at the next vertex
                                C.tsnap[2 * rec] = C.query_nee ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                                                               : make_float4(L.r, L.g, L.b, 0.0f);
                                C.tsnap[2 * rec + 1] = make_float4(beta.r, beta.g, beta.b, 0.0f);
                                C.tcount[tp] = nv + 1;
                            }
                        }
                    } else if (query_here) {
                        const int q = wave_alloc_active(&C.counters[0]);
                        cache_features(C, mat, sh.point, wo, sh.ns, C.qraw + 16 * size_t(q));
                        C.qbeta[q] = make_float4(beta.r, beta.g, beta.b, __int_as_float(pid));
                        if (C.query_nee && bounce < rc.max_bounces)
                            nee_only = true;
                        else
                            stop_here = true;
                    }
                }
                if (!stop_here && bounce++ < rc.max_bounces) {
                    const int32_t mi = S.slot_shade[slot].w;
                    if (mi < 0) {
                        // null material: continue straight (path_integrator.cpp:95-100)
                        --bounce;
                        new_o = sh.point;
                        new_d = -wo;
                        cont = true;
                    } else {
                        const bxdf b = make_bxdf(mat, sh.front, S);
                        const frame3 fr = frame_from_xz(sh.ts, sh.ns);
                        const f3 wo_l = to_local(fr, wo);
                        if (b.flags & (BX_DIFFUSE | BX_GLOSSY)) {
                            // SampleDirectLight: Next1D, Next2D drawn first
                            const float lu0 = pcg_float(st);
                            f2 lu12;
                            lu12.x = pcg_float(st);
                            lu12.y = pcg_float(st);
                            if (S.n_lights > 0) {
                                float pmf;
                                const int32_t li =
                                    rc.light_alias ? light_select_alias(S, lu0, pmf) : light_select(S, lu0, pmf, cdf);
                                const GpuLight& Lt = S.lights[li];
                                light_sample ls;
                                if (light_sample_li(Lt, sh.point, lu12, ls)) {
                                    const f3 wi_l = to_local(fr, ls.wi);
                                    const float bsdf_pdf = (wo_l.z == 0.0f) ? 0.0f : bxdf_pdf<KM>(b, wo_l, wi_l);
                                    if (!(is_black(ls.Li) || bsdf_pdf == 0.0f)) {
                                        const spec f = (wo_l.z == 0.0f) ? mks(0.0f) : bxdf_f<KM>(b, wo_l, wi_l);
                                        const float light_pdf = pmf * ls.pdf;
                                        const spec f_cos = f * absdot(sh.ns, ls.wi);
                                        if (Lt.type == BB_LIGHT_SPOT || Lt.type == BB_LIGHT_POINT ||
                                            Lt.type == BB_LIGHT_DIRECTIONAL) { // delta lights
                                            sh_c = ((beta * ls.Li) * f_cos) / light_pdf;
                                        } else {
                                            const float mis = power_heuristic(light_pdf, bsdf_pdf);
                                            sh_c = (((beta * mis) * ls.Li) * f_cos) / light_pdf;
                                        }
                                        shadow = true;
                                        sh_o = sh.point;
                                        sh_d = ls.wi;
                                        sh_tmax = ls.visibility;
                                    }
                                }
                            }
                        }
                        // Sample_f(wo, Next1D(), Next2D()): u12 drawn before u0
                        f2 u12;
                        u12.x = pcg_float(st);
                        u12.y = pcg_float(st);
                        const float u0 = pcg_float(st);
                        bsample s;
                        // at a query vertex with direct light the cache stands in for the continuation
                        bool ok = !nee_only && wo_l.z != 0.0f && (b.flags & (BX_REFLECTION | BX_TRANSMISSION)) &&
                                  bxdf_sample<KM>(b, wo_l, u0, u12, s);
                        if (ok) ok = !(is_black(s.f) || s.pdf == 0.0f || s.wi.z == 0.0f);
                        if (ok) {
                            const f3 wi = from_local(fr, s.wi);
                            specular = (s.flags & BX_SPECULAR) != 0;
                            if (s.flags & BX_TRANSMISSION) eta_scale *= sqr(s.eta);
                            prev_pdf = s.pdf;
                            beta = beta * ((s.f * absdot(sh.ns, wi)) / s.pdf);
                            new_o = sh.point;
                            new_d = wi;
                            cont = true;
                            // Russian roulette (path_integrator.cpp:133-146)
                            if (bounce > rc.rr_min_bounces) {
                                const float p = max_component(beta) * eta_scale;
                                if (p < 1.0f) {
                                    if (pcg_float(st) > p)
                                        cont = false;
                                    else
                                        beta = beta / p;
                                }
                            }
                        }
                    }
                }
            }
            P.L[pid] = make_float4(L.r, L.g, L.b, eta_scale);
            P.beta[pid] = make_float4(beta.r, beta.g, beta.b, prev_pdf);
            P.rng[pid] = st;
            P.flags[pid] = (bounce & 0xffff) | (specular ? kFlagSpecular : 0) | (fl & (kFlagTraining | (0xff << kStopShift)));
        }
        int ss = 0, ns = 0;
        block_alloc2(shadow, shadow_count, ss, cont, next_count, ns, s_alloc);
        if (shadow) {
            P.sray[2 * ss] = make_float4(sh_o.x, sh_o.y, sh_o.z, sh_tmax);
            P.sray[2 * ss + 1] = make_float4(sh_d.x, sh_d.y, sh_d.z, __int_as_float(pid));
            P.scontrib[ss] = make_float4(sh_c.r, sh_c.g, sh_c.b, 0.0f);
        }
        if (cont) {
            P.queue[out][ns] = pid;
            P.ray[out][2 * ns] = make_float4(new_o.x, new_o.y, new_o.z, 0.0f);
            P.ray[out][2 * ns + 1] = make_float4(new_d.x, new_d.y, new_d.z, 0.0f);
        }
    }
}

// the cache state at the start of a pass: queries only while the network has not diverged
__global__ void k_cache_pass_begin(CacheBufs C)
{
    if (threadIdx.x == 0) {
        C.state[0] = *C.diverged;
        C.state[1] = 0;
    }
}

// cached rendering: L += throughput x max(cache, 0) for every query.  A
// prediction that is not finite marks the network diverged and the pass
// (`pass`, counted from 0) for dropping: its query paths lost their tails.
__global__ __launch_bounds__(kShadeBlock) void k_cache_apply(CacheBufs C, float4* __restrict__ L, int pass)
{
    const int n = min(C.counters[0], C.max_queries);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 b = C.qbeta[i];
        const int pid = __float_as_int(b.w);
        const float* y = C.qpred + 3 * size_t(i);
        const float y0 = y[0], y1 = y[1], y2 = y[2];
        if (!(isfinite(y0) && isfinite(y1) && isfinite(y2))) {
            atomicOr(&C.state[1], 1);
            atomicOr(C.diverged, 1);
            atomicCAS(&C.state[2], 0, pass + 1);
            continue;
        }
        float4 l = L[pid];
        l.x = l.x + b.x * fmaxf(y0, 0.0f);
        l.y = l.y + b.y * fmaxf(y1, 0.0f);
        l.z = l.z + b.z * fmaxf(y2, 0.0f);
        L[pid] = l;
    }
}

// counts of the pass for the call's statistics: its queries and the rows its
// `steps` optimizer steps of `batch` rows trained on
__global__ void k_cache_tally(CacheBufs C, int steps, int batch)
{
    if (threadIdx.x != 0) return;
    C.tally[0] += (unsigned long long)min(max(C.counters[0], 0), C.max_queries);
    const int rows = min(max(C.counters[1], 0), C.max_rows);
    if (rows == 0 || batch <= 0) return;
    const int nb = (rows + batch - 1) / batch;
    unsigned long long t = 0;
    for (int s = 0; s < steps; ++s) t += (unsigned long long)min(batch, rows - (s % nb) * batch);
    C.tally[1] += t;
}

// training rows of the pass: at each recorded vertex of a training path the
// radiance it reflected towards the previous vertex, (L_final - L_vertex) /
// throughput_vertex per channel; vertices with a vanishing throughput or a
// non-finite target are skipped.  One thread per training path: it counts
// its rows, the block places all its rows with one atomic (an exclusive scan
// of the counts: same-address atomics per vertex were this kernel's limit),
// and the thread copies them as whole float4s.  Only the first `used` rows
// are written: the optimizer steps of the pass read no others (steps x batch
// rows, batch by batch from the start, when the pass makes more), and a
// block that finds the row counter at `used` already stops: the counter
// then still ends at `used` or more, which is all the steps' batching and
// the tally depend on.
__device__ __forceinline__ bool cache_target(const CacheBufs& C, float4 lf, size_t rec, float& tr, float& tg, float& tb)
{
    float4 ls = C.tsnap[2 * rec];
    const float4 bs = C.tsnap[2 * rec + 1];
    if (!(bs.x > 1e-6f && bs.y > 1e-6f && bs.z > 1e-6f)) return false;
    // indirect-only targets: a vertex the path did not continue from saw none
    if (C.query_nee && ls.w == 0.0f) ls = lf;
    tr = (lf.x - ls.x) / bs.x;
    tg = (lf.y - ls.y) / bs.y;
    tb = (lf.z - ls.z) / bs.z;
    return isfinite(tr) && isfinite(tg) && isfinite(tb);
}

__global__ __launch_bounds__(kShadeBlock) void k_cache_targets(CacheBufs C, const float4* __restrict__ L, int n_train,
                                                                int stride, int used)
{
    constexpr int kWaves = kShadeBlock / 64;
    __shared__ int s_wave[kWaves + 1];
    const int tp = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x == 0) s_wave[kWaves] = __atomic_load_n(&C.counters[1], __ATOMIC_RELAXED) >= used;
    __syncthreads();
    // no training once the network diverged; no rows any step would read
    if (C.state[0] || s_wave[kWaves]) return; // uniform over the block
    const bool on = tp < n_train;
    const float4 lf = on ? L[size_t(tp) * stride] : float4{ 0.0f, 0.0f, 0.0f, 0.0f };
    const int nv = on ? C.tcount[tp] : 0;
    int cnt = 0;
    for (int v = 0; v < nv; ++v) {
        float tr, tg, tb;
        cnt += cache_target(C, lf, size_t(tp) * C.max_vertices + v, tr, tg, tb) ? 1 : 0;
    }
    // inclusive scan of the counts over the wave, then over the waves
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    __syncthreads(); // s_wave[kWaves] read by every thread above
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kWaves; ++w) {
            const int x = s_wave[w];
            s_wave[w] = t;
            t += x;
        }
        s_wave[kWaves] = t ? atomicAdd(&C.counters[1], t) : 0;
    }
    __syncthreads();
    int r = s_wave[kWaves] + s_wave[wave] + inc - cnt;
    if (cnt == 0 || r >= used) return;
    for (int v = 0; v < nv && r < used; ++v) {
        const size_t rec = size_t(tp) * C.max_vertices + v;
        float tr, tg, tb;
        if (!cache_target(C, lf, rec, tr, tg, tb)) continue;
        const float4* src = reinterpret_cast<const float4*>(C.traw + 16 * rec);
        float4* dst = reinterpret_cast<float4*>(C.rows + 16 * size_t(r));
#pragma unroll
        for (int k = 0; k < 4; ++k) dst[k] = src[k];
        C.targets[3 * size_t(r)] = fmaxf(tr, 0.0f);
        C.targets[3 * size_t(r) + 1] = fmaxf(tg, 0.0f);
        C.targets[3 * size_t(r) + 2] = fmaxf(tb, 0.0f);
        ++r;
    }
}

// Film::AddSample in sample order (film.cpp:26-37); NaN/Inf samples dropped
// (integrator.cpp:193-196).  Path i of a pass is pixel i % npix, sample i / npix.
// skip (cached renders): non-zero drops the pass (it is rendered again without the cache)
__global__ __launch_bounds__(kShadeBlock) void k_film(RenderConsts rc, const float4* __restrict__ L, float* film_sum,
                                                       int32_t* film_count, const int32_t* skip)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= rc.npix || (skip && *skip)) return;
    float sr = film_sum[3 * p], sg = film_sum[3 * p + 1], sb = film_sum[3 * p + 2];
    int32_t cnt = film_count[p];
    for (int pid = p; pid < rc.n_paths; pid += rc.npix) {
        const float4 l = L[pid];
        if (!is_nullish(mks(l.x, l.y, l.z))) {
            sr = sr + l.x;
            sg = sg + l.y;
            sb = sb + l.z;
            cnt += 1;
        }
    }
    film_sum[3 * p] = sr;
    film_sum[3 * p + 1] = sg;
    film_sum[3 * p + 2] = sb;
    film_count[p] = cnt;
}

__global__ void k_zero_ints(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_set_int(int* p, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

} // namespace bb

// ============================================================ host control

namespace
{

using namespace bb;

struct Workspace
{
    size_t capacity = 0; // paths
    int max_depth_slots = 0;
    std::vector<void*> allocs;
    PathBufs P{};
    size_t gstack_entries = 0;
    void* gstack = nullptr;
    size_t stage_bytes = 0; // device copy of caller buffers that live in host memory
    void* stage = nullptr;
    uint64_t generation = 0; // bumped whenever a buffer above moves (captured graphs hold the pointers)
};

// A render captured as a HIP graph.  The key holds everything the launch
// sequence depends on except sample_begin, which k_raygen reads from the
// workspace: the camera, the parameters, the film pointers and the
// workspace generation.
struct GraphEntry
{
    std::vector<unsigned char> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
constexpr size_t kMaxGraphs = 8; // per scene, oldest dropped first

struct CacheWork
{
    size_t paths = 0, train_paths = 0;
    int max_vertices = 0;
    std::vector<void*> allocs;
    CacheBufs C{};
};

struct SceneExtra
{
    Workspace ws;
    CacheWork cache;
    float* filter_cdf = nullptr; // device copy of the Gaussian filter's CDFs
    float filter_sigma = 0.0f, filter_extent = 0.0f;
    std::vector<GraphEntry> graphs;
    hipStream_t capture = nullptr; // stream the graphs are captured on (default priority)
    // shadow rays of a depth traced beside the next depth's closest hits:
    // the second stream (default priority) and the two events that order it
    hipStream_t side = nullptr;
    hipEvent_t shaded = nullptr, shadowed = nullptr;
};

void drop_graph(GraphEntry& g)
{
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.graph) hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

std::vector<std::pair<BbScene*, SceneExtra*>> g_extras;
std::mutex g_extras_lock;

SceneExtra* extra_for(BbScene* s)
{
    std::lock_guard<std::mutex> g(g_extras_lock);
    for (auto& e : g_extras)
        if (e.first == s) return e.second;
    auto* x = new SceneExtra();
    g_extras.push_back({ s, x });
    return x;
}

template <typename T>
bool dalloc(Workspace& w, T** p, size_t count)
{
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(16, sizeof(T) * count)) != hipSuccess) return false;
    w.allocs.push_back(q);
    *p = reinterpret_cast<T*>(q);
    return true;
}

void free_workspace(Workspace& w)
{
    for (void* p : w.allocs) hipFree(p);
    w.allocs.clear();
    w.capacity = 0;
    w.max_depth_slots = 0;
    w.P = PathBufs{};
    if (w.gstack) hipFree(w.gstack);
    w.gstack = nullptr;
    w.gstack_entries = 0;
    if (w.stage) hipFree(w.stage);
    w.stage = nullptr;
    w.stage_bytes = 0;
}

bool ensure_stage(Workspace& w, size_t bytes)
{
    if (w.stage_bytes >= bytes) return true;
    if (w.stage) hipFree(w.stage);
    w.stage = nullptr;
    w.stage_bytes = 0;
    ++w.generation;
    if (hipMalloc(&w.stage, bytes) != hipSuccess) return false;
    w.stage_bytes = bytes;
    return true;
}

// true for device and managed memory, which kernels read directly; buffers in
// host memory (pageable or pinned) are staged through device memory instead
bool device_accessible(const void* p)
{
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError(); // unregistered host memory on some runtimes
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

// global overflow area of the traversal stack: (cap - lds_entries) entries
// for every thread of the persistent grid
bool ensure_gstack(Workspace& w, size_t entries)
{
    entries = std::max<size_t>(entries, 64);
    if (w.gstack_entries >= entries) {
        w.P.gstack = reinterpret_cast<int2*>(w.gstack);
        return true;
    }
    if (w.gstack) hipFree(w.gstack);
    w.gstack = nullptr;
    w.gstack_entries = 0;
    ++w.generation;
    if (hipMalloc(&w.gstack, entries * sizeof(int2)) != hipSuccess) return false;
    w.gstack_entries = entries;
    w.P.gstack = reinterpret_cast<int2*>(w.gstack);
    return true;
}

bool ensure_workspace(Workspace& w, size_t paths, int depth_slots)
{
    if (w.capacity >= paths && w.max_depth_slots >= depth_slots) return true;
    paths = std::max(paths, w.capacity);
    depth_slots = std::max(depth_slots, w.max_depth_slots);
    for (void* p : w.allocs) hipFree(p);
    w.allocs.clear();
    w.capacity = 0;
    w.max_depth_slots = 0;
    ++w.generation;
    PathBufs& P = w.P;
    bool ok = dalloc(w, &P.L, paths) && dalloc(w, &P.beta, paths) && dalloc(w, &P.rng, paths) &&
              dalloc(w, &P.flags, paths) && dalloc(w, &P.queue[0], paths) && dalloc(w, &P.queue[1], paths) &&
              dalloc(w, &P.ray[0], 2 * paths) && dalloc(w, &P.ray[1], 2 * paths) && dalloc(w, &P.hit, paths) &&
              dalloc(w, &P.sray, 2 * paths) && dalloc(w, &P.scontrib, paths) &&
              dalloc(w, &P.counts, size_t(4) * (depth_slots + 2)) && dalloc(w, &P.err, 4) && dalloc(w, &P.counters, 8) &&
              dalloc(w, &P.sample_begin, 1);
    if (!ok) {
        for (void* p : w.allocs) hipFree(p);
        w.allocs.clear();
        return false;
    }
    w.capacity = paths;
    w.max_depth_slots = depth_slots;
    w.P.gstack = reinterpret_cast<int2*>(w.gstack);
    return true;
}

// total memory of the current device in bytes (looked up once per device)
long long device_memory()
{
    static std::mutex lock;
    static long long bytes[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 16ll << 30;
    }
    std::lock_guard<std::mutex> g(lock);
    if (!bytes[dev]) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) {
            (void)hipGetLastError();
            total_b = size_t(16) << 30;
        }
        bytes[dev] = (long long)total_b;
    }
    return bytes[dev];
}

int grid_for(long long n, int block)
{
    long long g = (n + block - 1) / block;
    return int(std::max<long long>(1, std::min<long long>(g, kMaxGridBlocks)));
}

// PerspectiveCamera constructor (src/camera/perspective_camera.cpp:8-38)
GpuCamera make_camera(const BbCameraDesc& c)
{
    GpuCamera g;
    quat q{ c.orientation[0], c.orientation[1], c.orientation[2], c.orientation[3] };
    float theta = c.vfov_deg * kPi / 180.0f;
    float h = std::tan(theta / 2.0f);
    float aspect_ratio = float(c.width) / float(c.height);
    float viewport_height = 2.0f * h;
    float viewport_width = aspect_ratio * viewport_height;
    // Quat::GetBasisX/Y/Z (quaternion.h:181-205)
    float x2 = q.x * 2.0f, y2 = q.y * 2.0f, z2 = q.z * 2.0f, w2 = q.w * 2.0f;
    f3 bx = mk3((q.w * w2) - 1.0f + q.x * x2, (q.z * w2) + q.y * x2, (-q.y * w2) + q.z * x2);
    f3 by = mk3((-q.z * w2) + q.x * y2, (q.w * w2) - 1.0f + q.y * y2, (q.x * w2) + q.z * y2);
    f3 bz = mk3((q.y * w2) + q.x * z2, (-q.x * w2) + q.y * z2, (q.w * w2) - 1.0f + q.z * z2);
    g.u = -bx;
    g.v = by;
    f3 w = bz;
    g.origin = mk3(c.position[0], c.position[1], c.position[2]);
    g.horizontal = g.u * (c.focus_distance * viewport_width);
    g.vertical = g.v * (c.focus_distance * viewport_height);
    g.lower_left = ((g.origin - g.horizontal / 2.0f) - g.vertical / 2.0f) + w * c.focus_distance;
    g.lens_radius = c.aperture_radius;
    g.filter_extent = c.filter_extent;
    g.width = c.width;
    g.height = c.height;
    g.filter_type = c.filter_type;
    g.filter_cdf = nullptr;
    return g;
}

// GaussianFilter::GaussianFilter (gaussian_filter.cpp:8-27) and the CDFs of
// its Distribution2D (sampling.h:343-379, 426-445), computed on the host:
// out = marginal CDF, then the 32 conditional CDFs, 33 floats each
void gaussian_filter_cdfs(float sigma, float extent, float* out)
{
    const int n = kFilterN;
    auto gaussian = [](float x, float sg) {
        // Gaussian(x, 0, sigma) (sampling.h:9-12)
        const float xm = x - 0.0f;
        return 1.0f / std::sqrt(2.0f * kPi * sg * sg) * std::exp(-(xm * xm) / (2.0f * sg * sg));
    };
    auto dist1d = [n](const float* f, float* cdf) {
        cdf[0] = 0.0f;
        for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / float(n);
        const float integral = cdf[n];
        if (integral == 0.0f)
            for (int i = 1; i < n + 1; ++i) cdf[i] = float(i) / float(n);
        else
            for (int i = 1; i < n + 1; ++i) cdf[i] /= integral;
        return integral;
    };
    std::vector<float> values(size_t(n) * n), marginal(n);
    const float half_extent = extent / 2.0f;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float px = ((float(x) + 0.5f) / float(n)) * extent - half_extent;
            const float py = ((float(y) + 0.5f) / float(n)) * extent - half_extent;
            values[size_t(y) * n + x] = gaussian(px, sigma) * gaussian(py, sigma);
        }
    for (int v = 0; v < n; ++v) marginal[v] = dist1d(&values[size_t(v) * n], out + (n + 1) * (1 + v));
    dist1d(marginal.data(), out);
}

#define HIP_OK(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) {                                                                                        \
            bb::set_error("%s: %s", #call, hipGetErrorString(e_));                                                     \
            return BBGPU_ERR_DEVICE;                                                                                   \
        }                                                                                                              \
    } while (0)

struct EventPool
{
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    hipEvent_t get()
    {
        if (used == ev.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            ev.push_back(e);
        }
        return ev[used++];
    }
    ~EventPool()
    {
        for (auto e : ev) hipEventDestroy(e);
    }
};

// LDS entries of the traversal stack per lane: `requested` or, by default,
// the whole stack up to 32 entries (deeper entries overflow to global memory)
int lds_entries_for(const SceneView& V, int requested, bool count)
{
    // the counting kernels (measurement only) keep their whole one-level stack in LDS
    int e = requested > 0 ? requested : (count ? 32 : kLdsEntriesDefault);
    return std::max(1, std::min(e, trace_stack_depth(V, count)));
}

// dynamic LDS of a traversal kernel: 4-byte entries for any-hit queries
// without counters, 8-byte ones otherwise
size_t trace_lds_bytes(int lds_entries, bool slim = false, int block = kTraceBlock)
{
    return size_t(lds_entries) * size_t(block) * (slim ? sizeof(int) : sizeof(int2));
}

// persistent grid of a traversal kernel: resident blocks per CU x CUs
int trace_grid(const void* kernel, size_t lds_bytes, int block = kTraceBlock)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds_bytes) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    return per_cu * cus;
}

// Shading kernel for a scene: the texture code only for scenes that use
// textures, and the BxDF code of the scene's material kinds only (two
// common sets: Lambertian + conductor + dielectric, and metallic-roughness +
// Lambertian), so the common kernels need fewer registers
constexpr int kKindsLcd = (1 << MK_LAMBERT) | (1 << MK_CONDUCTOR) | (1 << MK_DIELECTRIC);
constexpr int kKindsMr = (1 << MK_METALLIC_ROUGHNESS) | (1 << MK_LAMBERT);
using ShadeKernel = void (*)(SceneView, RenderConsts, PathBufs, int, CacheBufs);

template <bool CACHE>
ShadeKernel shade_kernel_t(const SceneView& V)
{
    if (V.has_textures) return k_shade<CACHE, true, kAllKinds>;
    if ((V.kind_mask & ~kKindsLcd) == 0) return k_shade<CACHE, false, kKindsLcd>;
    if ((V.kind_mask & ~kKindsMr) == 0) return k_shade<CACHE, false, kKindsMr>;
    return k_shade<CACHE, false, kAllKinds>;
}

ShadeKernel shade_kernel(const SceneView& V, bool cache)
{
    return cache ? shade_kernel_t<true>(V) : shade_kernel_t<false>(V);
}

using BatchKernel = void (*)(SceneView, BatchSrc, int, int*, int2*, int, int, unsigned long long*, int*);
using ClosestKernel = void (*)(SceneView, const int*, int*, ClosestSrc, int2*, int, int, unsigned long long*, int*);
using ShadowKernel = void (*)(SceneView, const int*, int*, ShadowSrc, int2*, int, int, unsigned long long*, int*);

// the traversal variants: any-hit or closest, counters, whole stack in LDS,
// and the alpha test compiled in only for scenes that have alpha textures
template <bool ANY, bool ALPHA>
BatchKernel batch_kernel_t(bool count, bool full)
{
    if (count)
        return full ? k_trace_batch<ANY, true, kStackLds, ALPHA> : k_trace_batch<ANY, true, kStackSplit, ALPHA>;
    return full ? k_trace_batch<ANY, false, kStackLds, ALPHA> : k_trace_batch<ANY, false, kStackSplit, ALPHA>;
}

BatchKernel batch_kernel(bool any, bool count, bool full, bool alpha)
{
    if (alpha) return any ? batch_kernel_t<true, true>(count, full) : batch_kernel_t<false, true>(count, full);
    return any ? batch_kernel_t<true, false>(count, full) : batch_kernel_t<false, false>(count, full);
}

// The bounce loop's traversal variants: counters (measurement only: 128-thread
// blocks, stack in LDS), stack placement SK, the alpha test compiled in only
// for scenes that have alpha textures, and the block size (64, 128 or 256)
template <bool COUNT, bool ALPHA>
ClosestKernel closest_sel(int sk, int block)
{
    if constexpr (COUNT)
        return sk == kStackLds ? k_trace_closest<true, kStackLds, ALPHA, kTraceBlock>
                               : k_trace_closest<true, kStackSplit, ALPHA, kTraceBlock>;
    else {
        if (sk == kStackScratch)
            return block == 64 ? k_trace_closest<false, kStackScratch, ALPHA, 64>
                               : block == 256 ? k_trace_closest<false, kStackScratch, ALPHA, 256>
                                              : k_trace_closest<false, kStackScratch, ALPHA, kTraceBlock>;
        if (sk == kStackLds)
            return block == 64 ? k_trace_closest<false, kStackLds, ALPHA, 64>
                               : block == 256 ? k_trace_closest<false, kStackLds, ALPHA, 256>
                                              : k_trace_closest<false, kStackLds, ALPHA, kTraceBlock>;
        return block == 64 ? k_trace_closest<false, kStackSplit, ALPHA, 64>
                           : block == 256 ? k_trace_closest<false, kStackSplit, ALPHA, 256>
                                          : k_trace_closest<false, kStackSplit, ALPHA, kTraceBlock>;
    }
}

ClosestKernel closest_kernel(bool count, int sk, bool alpha, int block)
{
    if (alpha) return count ? closest_sel<true, true>(sk, block) : closest_sel<false, true>(sk, block);
    return count ? closest_sel<true, false>(sk, block) : closest_sel<false, false>(sk, block);
}

template <bool COUNT, bool ALPHA>
ShadowKernel shadow_sel(int sk, int block)
{
    if constexpr (COUNT)
        return sk == kStackLds ? k_trace_shadow<true, kStackLds, ALPHA, kTraceBlock>
                               : k_trace_shadow<true, kStackSplit, ALPHA, kTraceBlock>;
    else {
        if (sk == kStackScratch)
            return block == 64 ? k_trace_shadow<false, kStackScratch, ALPHA, 64>
                               : block == 256 ? k_trace_shadow<false, kStackScratch, ALPHA, 256>
                                              : k_trace_shadow<false, kStackScratch, ALPHA, kTraceBlock>;
        if (sk == kStackLds)
            return block == 64 ? k_trace_shadow<false, kStackLds, ALPHA, 64>
                               : block == 256 ? k_trace_shadow<false, kStackLds, ALPHA, 256>
                                              : k_trace_shadow<false, kStackLds, ALPHA, kTraceBlock>;
        return block == 64 ? k_trace_shadow<false, kStackSplit, ALPHA, 64>
                           : block == 256 ? k_trace_shadow<false, kStackSplit, ALPHA, 256>
                                          : k_trace_shadow<false, kStackSplit, ALPHA, kTraceBlock>;
    }
}

ShadowKernel shadow_kernel(bool count, int sk, bool alpha, int block)
{
    if (alpha) return count ? shadow_sel<true, true>(sk, block) : shadow_sel<false, true>(sk, block);
    return count ? shadow_sel<true, false>(sk, block) : shadow_sel<false, false>(sk, block);
}

} // namespace

namespace
{

// What a traversal batch launches: the kernel variant, its LDS, the number
// of stack entries kept in LDS and the persistent grid
struct TracePlan
{
    BatchKernel kfn;
    size_t lds;
    int lds_entries;
    int grid;
    size_t overflow_per_lane; // stack entries beyond LDS per lane of the grid
};

TracePlan trace_plan(const SceneView& V, int32_t n_rays, int32_t any_hit, bool count)
{
    TracePlan t{};
    const int depth = trace_stack_depth(V, count);
    t.lds_entries = lds_entries_for(V, 0, count);
    t.lds = trace_lds_bytes(t.lds_entries, any_hit != 0 && !count);
    const bool full = depth <= t.lds_entries;
    t.kfn = batch_kernel(any_hit != 0, count, full, V.has_alpha != 0);
    t.grid = std::min(trace_grid(reinterpret_cast<const void*>(t.kfn), t.lds), grid_for(std::max(n_rays, 1), kTraceBlock));
    t.overflow_per_lane = size_t(std::max(0, depth - t.lds_entries));
    return t;
}

// The device memory a traversal batch uses beyond its rays and hits
struct TraceScratch
{
    int* err;                     // 1: a traversal stack overflowed
    int* counts;                  // [4]: the work-queue fetch counter at [0]
    unsigned long long* counters; // [8]: nodes, prims, accepts
    int2* gstack;                 // the stack's overflow area, plan.overflow_per_lane entries per lane
};

// Traces a batch whose rays and hits are in device memory; returns once it is done
int trace_run(const char* who, const SceneView& V, const BbRay* rays, int32_t n_rays, int32_t any_hit, BbHit* hits,
              BbTraceCounters* counters, hipStream_t st, const TracePlan& t, const TraceScratch& s)
{
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, s.err, 1);
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, s.counts, 4);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, s.counters, 8);
    if (n_rays > 0) {
        BatchSrc src{ reinterpret_cast<const float4*>(rays), reinterpret_cast<int4*>(hits), V.tris, any_hit };
        hipLaunchKernelGGL(t.kfn, dim3(t.grid), dim3(kTraceBlock), t.lds, st, V, src, n_rays, s.counts, s.gstack,
                           t.lds_entries, 48, s.counters, s.err);
        HIP_OK(hipGetLastError());
    }
    int err = 0;
    unsigned long long c[8] = { 0 };
    HIP_OK(hipMemcpyAsync(&err, s.err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(c, s.counters, sizeof(c), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (err) {
        bb::set_error("%s: traversal stack overflow", who);
        return BBGPU_ERR_DEVICE;
    }
    if (counters) {
        counters->rays = uint64_t(n_rays);
        counters->nodes = c[0];
        counters->prims = c[1];
        counters->accepts = c[2];
    }
    return BBGPU_OK;
}

// Scratch of bbgpu_trace_noalloc, one copy per device (it lives as long as
// the process): a 256-byte header (counters at 0, the error flag at 64, the
// work-queue counts at 128), then the traversal stacks' overflow area.  Host
// code never passes `scratch` itself to a kernel (its host-side name is not a
// device address); it uses the address hipGetSymbolAddress returns for the
// current device.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr size_t kNoAllocHeaderBytes = 256;

} // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace
{

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per device)
unsigned char* scratch_of_current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

} // namespace

extern "C" {

int bbgpu_trace(BbScene* scene, const BbRay* rays_dev, int32_t n_rays, int32_t any_hit, BbHit* hits_dev,
                BbTraceCounters* counters, void* stream)
{
    if (!scene || n_rays < 0 || (n_rays > 0 && (!rays_dev || !hits_dev))) {
        bb::set_error("bbgpu_trace: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    SceneExtra* x = extra_for(scene);
    const SceneView& V = scene->view;
    const TracePlan t = trace_plan(V, n_rays, any_hit, counters != nullptr);
    if (!ensure_workspace(x->ws, 1, 16) || !ensure_gstack(x->ws, t.overflow_per_lane * size_t(t.grid) * kTraceBlock)) {
        bb::set_error("bbgpu_trace: workspace allocation failed");
        return BBGPU_ERR_NOMEM;
    }
    PathBufs& P = x->ws.P;
    const BbRay* rays = rays_dev;
    BbHit* hits = hits_dev;
    const bool host_rays = n_rays > 0 && !device_accessible(rays_dev);
    const bool host_hits = n_rays > 0 && !device_accessible(hits_dev);
    if (host_rays || host_hits) {
        const size_t rb = host_rays ? sizeof(BbRay) * size_t(n_rays) : 0;
        const size_t hb = host_hits ? sizeof(BbHit) * size_t(n_rays) : 0;
        if (!ensure_stage(x->ws, rb + hb)) {
            bb::set_error("bbgpu_trace: staging allocation failed");
            return BBGPU_ERR_NOMEM;
        }
        char* stage = reinterpret_cast<char*>(x->ws.stage);
        if (host_rays) {
            rays = reinterpret_cast<const BbRay*>(stage);
            HIP_OK(hipMemcpyAsync(stage, rays_dev, rb, hipMemcpyHostToDevice, st));
        }
        if (host_hits) hits = reinterpret_cast<BbHit*>(stage + rb);
    }
    const TraceScratch s{ P.err, P.counts, P.counters, P.gstack };
    const int r = trace_run("bbgpu_trace", V, rays, n_rays, any_hit, hits, counters, st, t, s);
    if (r == BBGPU_OK && host_hits)
        HIP_OK(hipMemcpy(hits_dev, hits, sizeof(BbHit) * size_t(n_rays), hipMemcpyDeviceToHost));
    return r;
}

int bbgpu_trace_noalloc(BbScene* scene, const BbRay* rays_dev, int32_t n_rays, int32_t any_hit, BbHit* hits_dev,
                        BbTraceCounters* counters, void* stream)
{
    if (!scene || n_rays < 0 || (n_rays > 0 && (!rays_dev || !hits_dev))) {
        bb::set_error("bbgpu_trace_noalloc: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    if (n_rays > 0 && (!device_accessible(rays_dev) || !device_accessible(hits_dev))) {
        bb::set_error("bbgpu_trace_noalloc: rays and hits must be in device memory");
        return BBGPU_ERR_INVALID;
    }
    unsigned char* base = scratch_of_current_device();
    if (!base) {
        bb::set_error("bbgpu_trace_noalloc: the scratch array has no address on the current device");
        return BBGPU_ERR_DEVICE;
    }
    const SceneView& V = scene->view;
    TracePlan t = trace_plan(V, n_rays, any_hit, counters != nullptr);
    // a smaller persistent grid when the stacks' overflow area would not fit
    const size_t area = kNoAllocScratchBytes - kNoAllocHeaderBytes;
    if (t.overflow_per_lane > 0) {
        const size_t per_block = t.overflow_per_lane * size_t(kTraceBlock) * sizeof(int2);
        const size_t fit = area / per_block;
        if (fit < 1) {
            bb::set_error("bbgpu_trace_noalloc: the traversal stack (%d entries, %d in LDS) needs %zu bytes of "
                          "scratch per block, more than the %zu bytes available",
                          trace_stack_depth(V, counters != nullptr), t.lds_entries, per_block, area);
            return BBGPU_ERR_INVALID;
        }
        t.grid = int(std::min<size_t>(size_t(t.grid), fit));
    }
    const TraceScratch s{ reinterpret_cast<int*>(base + 64), reinterpret_cast<int*>(base + 128),
                          reinterpret_cast<unsigned long long*>(base),
                          reinterpret_cast<int2*>(base + kNoAllocHeaderBytes) };
    return trace_run("bbgpu_trace_noalloc", V, rays_dev, n_rays, any_hit, hits_dev, counters, (hipStream_t)stream, t, s);
}

} // extern "C"

namespace bb
{
int nrc_query_counted(BbNrc* x, const float* inputs_dev, const int* count_dev, int cap, float* out_dev, hipStream_t st);
int nrc_train_counted(BbNrc* x, const float* rows_dev, const float* targets_dev, const int* count_dev, int cap,
                      int batch, int step, hipStream_t st);
int* nrc_status(BbNrc* x);
bool nrc_diverged(const BbNrc* x);
void nrc_note_diverged(BbNrc* x, bool d);
float* nrc_loss(BbNrc* x);
} // namespace bb

namespace
{

// the radiance cache of a cached render
struct CacheCtx
{
    BbNrc* nrc;
    BbCacheParams cp;
    BbCacheStats* stats;
};

bool ensure_cache(SceneExtra* x, size_t cap, int stride, int max_vertices)
{
    CacheWork& w = x->cache;
    const size_t ntp = (cap + size_t(stride) - 1) / size_t(stride);
    if (w.paths >= cap && w.train_paths >= ntp && w.max_vertices >= max_vertices) return true;
    for (void* q : w.allocs) (void)hipFree(q);
    w.allocs.clear();
    w.paths = w.train_paths = 0;
    w.max_vertices = 0;
    auto get = [&](void** q, size_t bytes) {
        if (hipMalloc(q, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
        w.allocs.push_back(*q);
        return true;
    };
    CacheBufs& C = w.C;
    const size_t rows = ntp * size_t(max_vertices);
    const bool ok = get((void**)&C.qraw, cap * 64) && get((void**)&C.qbeta, cap * 16) && get((void**)&C.qpred, cap * 12) &&
                    get((void**)&C.traw, rows * 64) && get((void**)&C.tsnap, rows * 32) &&
                    get((void**)&C.tcount, ntp * 4) && get((void**)&C.rows, rows * 64) &&
                    get((void**)&C.targets, rows * 12) && get((void**)&C.counters, 16) &&
                    get((void**)&C.state, 16) && get((void**)&C.tally, 16);
    if (!ok) {
        for (void* q : w.allocs) (void)hipFree(q);
        w.allocs.clear();
        return false;
    }
    w.paths = cap;
    w.train_paths = ntp;
    w.max_vertices = max_vertices;
    C.max_vertices = max_vertices;
    C.max_rows = int32_t(rows);
    C.max_queries = int32_t(cap);
    return true;
}

// wait = false (batched jobs): return once every launch is queued; the
// caller synchronizes the stream and checks the error flag (check_render)
// scenes with fewer primitives trace shadow rays on the render stream
constexpr int32_t kOverlapMinPrims = 8;

// `overlap`: trace each depth's shadow rays on a second stream, beside the
// next depth's closest hits (batches pass false: their jobs overlap already)
int render_impl(BbScene* scene, const BbCameraDesc* camera, const BbRenderParams* params, float* film_sum_dev,
                int32_t* film_count_dev, void* stream, BbRenderStats* stats, int32_t count_traversal, CacheCtx* cache,
                bool wait = true, bool overlap = true)
{
    if (!scene || !camera || !params || !film_sum_dev || !film_count_dev) {
        bb::set_error("bbgpu_render: null argument");
        return BBGPU_ERR_INVALID;
    }
    const BbRenderParams& p = *params;
    if (camera->width <= 0 || camera->height <= 0 || p.spp < 0 || p.max_bounces < 0 || p.sample_begin < 0) {
        bb::set_error("bbgpu_render: invalid camera or parameters");
        return BBGPU_ERR_INVALID;
    }
    if (camera->filter_type < BB_FILTER_BOX || camera->filter_type > BB_FILTER_GAUSSIAN ||
        !(camera->filter_extent > 0.0f) || !std::isfinite(camera->filter_extent) ||
        (camera->filter_type == BB_FILTER_GAUSSIAN && !(camera->filter_sigma > 0.0f))) {
        bb::set_error("bbgpu_render: invalid filter (type %d, extent %g, sigma %g)", camera->filter_type,
                      double(camera->filter_extent), double(camera->filter_sigma));
        return BBGPU_ERR_INVALID;
    }
    if (p.regularize_bsdf) {
        bb::set_error("bbgpu_render: regularize_bsdf is not supported");
        return BBGPU_ERR_UNSUPPORTED;
    }
    if (p.light_sampler != BB_LIGHTS_POWER && p.light_sampler != BB_LIGHTS_ALIAS) {
        bb::set_error("bbgpu_render: unknown light sampler %d", p.light_sampler);
        return BBGPU_ERR_INVALID;
    }
    if (scene->view.has_null_material) {
        bb::set_error("bbgpu_render: primitives without material are not supported");
        return BBGPU_ERR_UNSUPPORTED;
    }
    if (p.max_bounces > 1000) {
        bb::set_error("bbgpu_render: max_bounces above 1000");
        return BBGPU_ERR_INVALID;
    }
    const long long npix = (long long)camera->width * camera->height;
    if (npix > (1ll << 30)) {
        bb::set_error("bbgpu_render: image too large");
        return BBGPU_ERR_INVALID;
    }
    if (p.spp == 0) return BBGPU_OK;
    // Paths in flight per pass.  Deep bounces hold few rays, and a traversal
    // launch lasts as long as its slowest ray, so fewer, fuller launches win:
    // by default a pass takes as many samples as an eighth of the device's
    // memory holds in path state (a whole 1080p frame of 64 spp is 133M
    // paths, about 24 GB)
    long long want = p.wavefront_paths;
    if (want <= 0) {
        const long long bytes_per_path = cache ? 384 : 192;
        want = std::max<long long>(1ll << 22, std::min<long long>(1ll << 28, device_memory() / 8 / bytes_per_path));
    }
    long long per_pass = std::max<long long>(1, want / npix); // samples per pass
    per_pass = std::min<long long>(per_pass, p.spp);
    const long long cap = per_pass * npix;
    if (cap > (1ll << 30)) {
        bb::set_error("bbgpu_render: wavefront too large");
        return BBGPU_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    SceneExtra* x = extra_for(scene);
    if (!ensure_workspace(x->ws, size_t(cap), p.max_bounces + 1)) {
        bb::set_error("bbgpu_render: workspace allocation of %lld paths failed", cap);
        return BBGPU_ERR_NOMEM;
    }
    const SceneView& V = scene->view;
    const bool timed = stats != nullptr;
    const bool count = timed && count_traversal;
    const int depth = trace_stack_depth(V, count);
    // traversal launch shape: block size, and where the stacks live
    const int block = p.trace_block > 0 ? p.trace_block : kTraceBlock;
    if (block != 64 && block != 128 && block != 256) {
        bb::set_error("bbgpu_render: trace_block %d (64, 128 or 256)", p.trace_block);
        return BBGPU_ERR_INVALID;
    }
    if (p.trace_stack != BB_STACK_LDS && p.trace_stack != BB_STACK_SCRATCH) {
        bb::set_error("bbgpu_render: unknown trace_stack %d", p.trace_stack);
        return BBGPU_ERR_INVALID;
    }
    const bool scratch_stack = p.trace_stack == BB_STACK_SCRATCH && !count;
    if (scratch_stack && depth > kScratchStackEntries) {
        bb::set_error("bbgpu_render: the scratch stack holds %d entries, this scene's traversal needs %d",
                      kScratchStackEntries, depth);
        return BBGPU_ERR_UNSUPPORTED;
    }
    const int tblock = count ? kTraceBlock : block; // the counting kernels run 128-thread blocks
    const int lds_entries = scratch_stack ? 0 : lds_entries_for(V, p.trace_lds_stack, count);
    const size_t lds = scratch_stack ? 0 : trace_lds_bytes(lds_entries, false, tblock);
    // reference-only entries without counters
    const size_t lds_shadow = scratch_stack ? 0 : trace_lds_bytes(lds_entries, !count, tblock);
    const int refill_min = p.trace_refill > 0 ? std::min(p.trace_refill, 64) : kTraceRefillDefault;
    const int refill_shadow = p.shadow_refill > 0 ? std::min(p.shadow_refill, 64) : kShadowRefillDefault;
    const int sk = scratch_stack ? kStackScratch : (depth <= lds_entries ? kStackLds : kStackSplit);
    const ClosestKernel k_closest = closest_kernel(count, sk, V.has_alpha != 0, tblock);
    // shading with the texture code only for scenes that use textures
    void (*const k_shade_fn)(SceneView, RenderConsts, PathBufs, int, CacheBufs) = shade_kernel(V, cache != nullptr);
    const ShadowKernel k_shadow = shadow_kernel(count, sk, V.has_alpha != 0, tblock);
    const int g_trace =
        std::min(trace_grid(reinterpret_cast<const void*>(k_closest), lds, tblock), grid_for(cap, tblock));
    const int g_shadow =
        std::min(trace_grid(reinterpret_cast<const void*>(k_shadow), lds_shadow, tblock), grid_for(cap, tblock));
    // the overflow area serves both grids (int2-sized entries hold the 4-byte ones too)
    const size_t overflow = sk == kStackSplit ? size_t(depth - lds_entries) : 0;
    // Shadow rays of depth d and closest hits of depth d + 1 read only what
    // the shading of depth d wrote, so they may run at once; the shading of
    // depth d + 1 waits for both (it adds to the radiance the shadow rays
    // add to, and rewrites the shadow-ray queue), so every sum keeps its
    // order and the film is bit-identical.  The two kernels then take
    // separate halves of the overflow area.  A captured graph records the
    // same fork and join.  Not for counting frames, which keep one stream,
    // nor for scenes so small that the two handoffs per depth (some
    // microseconds each) cost more than the traversal tails they hide: a
    // one-sphere 1080p frame measured 2.00 ms on one stream and 2.15 ms on
    // two, a sphere on a floor 2.51 and 2.55, a 13-primitive closed box 3.06
    // and 3.02, a 338-triangle displaced icosphere 3.77 and 3.17.
    const bool split = overlap && !count && scene->n_prims >= kOverlapMinPrims;
    const size_t ov_half = overflow * size_t(std::max(g_trace, g_shadow)) * size_t(tblock);
    if (!ensure_gstack(x->ws, ov_half * (split ? 2 : 1))) {
        bb::set_error("bbgpu_render: traversal stack allocation failed");
        return BBGPU_ERR_NOMEM;
    }
    if (split && (!x->side || !x->shaded || !x->shadowed)) {
        bool ok = true;
        if (!x->shaded && hipEventCreateWithFlags(&x->shaded, hipEventDisableTiming) != hipSuccess) {
            x->shaded = nullptr;
            ok = false;
        }
        if (ok && !x->shadowed && hipEventCreateWithFlags(&x->shadowed, hipEventDisableTiming) != hipSuccess) {
            x->shadowed = nullptr;
            ok = false;
        }
        if (ok && !x->side && hipStreamCreateWithFlags(&x->side, hipStreamNonBlocking) != hipSuccess) { // default priority
            x->side = nullptr;
            ok = false;
        }
        if (!ok) {
            bb::set_error("bbgpu_render: shadow stream creation failed");
            return BBGPU_ERR_DEVICE;
        }
    }
    PathBufs& P = x->ws.P;
    GpuCamera cam = make_camera(*camera);
    if (camera->filter_type == BB_FILTER_GAUSSIAN) {
        if (!x->filter_cdf || x->filter_sigma != camera->filter_sigma || x->filter_extent != camera->filter_extent) {
            std::vector<float> cdf(size_t(kFilterN + 1) * (kFilterN + 1));
            gaussian_filter_cdfs(camera->filter_sigma, camera->filter_extent, cdf.data());
            if (!x->filter_cdf && hipMalloc(&x->filter_cdf, sizeof(float) * cdf.size()) != hipSuccess) {
                x->filter_cdf = nullptr;
                bb::set_error("bbgpu_render: filter table allocation failed");
                return BBGPU_ERR_NOMEM;
            }
            HIP_OK(hipMemcpyAsync(x->filter_cdf, cdf.data(), sizeof(float) * cdf.size(), hipMemcpyHostToDevice, st));
            HIP_OK(hipStreamSynchronize(st));
            x->filter_sigma = camera->filter_sigma;
            x->filter_extent = camera->filter_extent;
        }
        cam.filter_cdf = x->filter_cdf;
    }
    // a film in host memory is copied in, accumulated on the device and copied back
    float* const film_sum_user = film_sum_dev;
    int32_t* const film_count_user = film_count_dev;
    const bool host_film = !device_accessible(film_sum_dev) || !device_accessible(film_count_dev);
    const size_t sum_bytes = sizeof(float) * 3 * size_t(npix), count_bytes = sizeof(int32_t) * size_t(npix);
    if (host_film) {
        if (!ensure_stage(x->ws, sum_bytes + count_bytes)) {
            bb::set_error("bbgpu_render: staging allocation failed");
            return BBGPU_ERR_NOMEM;
        }
        film_sum_dev = reinterpret_cast<float*>(x->ws.stage);
        film_count_dev = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(x->ws.stage) + sum_bytes);
        HIP_OK(hipMemcpyAsync(film_sum_dev, film_sum_user, sum_bytes, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(film_count_dev, film_count_user, count_bytes, hipMemcpyHostToDevice, st));
    }
    if (timed) memset(stats, 0, sizeof(*stats));
    const int stride = cache ? std::max(1, cache->cp.train_stride) : 0;
    CacheBufs* C = nullptr;
    if (cache) {
        if (!ensure_cache(x, size_t(cap), stride, p.max_bounces + 1)) {
            bb::set_error("bbgpu_render_cached: cache workspace allocation failed");
            return BBGPU_ERR_NOMEM;
        }
        C = &x->cache.C;
        const f3 lo = mk3(V.root_box[0], V.root_box[1], V.root_box[2]);
        const f3 hi = mk3(V.root_box[3], V.root_box[4], V.root_box[5]);
        const f3 ext = hi - lo;
        C->wmin = lo;
        C->winv = mk3(1.0f / std::max(ext.x, 1e-6f), 1.0f / std::max(ext.y, 1e-6f), 1.0f / std::max(ext.z, 1e-6f));
        C->query_bounce = std::max(1, cache->cp.query_bounce);
        C->query_first = cache->cp.query_primary ? 0 : std::max(1, cache->cp.query_first);
        C->query_nee = cache->cp.query_direct ? 1 : 0;
        C->extra = std::max(0, cache->cp.train_extra_bounces);
        C->diverged = nrc_status(cache->nrc);
        if (cache->stats) memset(cache->stats, 0, sizeof(*cache->stats));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, C->tally, 2);
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, C->state, 4);
    }
    CacheBufs Cv{};
    if (C) Cv = *C;

    EventPool pool;
    struct Span
    {
        hipEvent_t a, b;
        int kind; // 0 closest, 1 shadow, 2 shade, 3 cache query, 4 cache training
    };
    std::vector<Span> spans;
    const bool cache_timed = cache && cache->stats;
    auto mark = [&]() { hipEvent_t e = pool.get(); hipEventRecord(e, st); return e; };
    auto mark_on = [&](hipStream_t s) { hipEvent_t e = pool.get(); hipEventRecord(e, s); return e; };

    hipEvent_t t_begin = nullptr;
    if (timed) t_begin = mark();
    hipEvent_t t_begin_all = (cache && cache->stats) ? mark() : nullptr;
    // the one value that differs between calls that share a captured graph
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, P.sample_begin, p.sample_begin);

    // Graph mode: the launch sequence below has no host synchronization when
    // neither stats nor the cache are on; it is captured once on the scene's
    // capture stream and replayed on the caller's stream
    const bool use_graph = p.graph != 0 && !timed && !cache;
    GraphEntry* graph = nullptr;
    bool capturing = false;
    std::vector<unsigned char> key;
    hipStream_t ls = st; // the stream the launch sequence goes to
    if (use_graph) {
        BbRenderParams kp = p;
        kp.sample_begin = 0;
        const void* films[2] = { film_sum_dev, film_count_dev };
        const uint64_t gen = x->ws.generation;
        auto put = [&](const void* d, size_t n) {
            key.insert(key.end(), static_cast<const unsigned char*>(d), static_cast<const unsigned char*>(d) + n);
        };
        put(camera, sizeof(*camera));
        put(&kp, sizeof(kp));
        put(films, sizeof(films));
        put(&gen, sizeof(gen));
        for (GraphEntry& ge : x->graphs)
            if (ge.key == key) graph = &ge;
        if (!graph) {
            if (!x->capture && hipStreamCreateWithFlags(&x->capture, hipStreamNonBlocking) != hipSuccess) {
                x->capture = nullptr;
                bb::set_error("bbgpu_render: capture stream creation failed");
                return BBGPU_ERR_DEVICE;
            }
            HIP_OK(hipStreamBeginCapture(x->capture, hipStreamCaptureModeThreadLocal));
            ls = x->capture;
            capturing = true;
        }
    }
    const int nslots = 4 * (p.max_bounces + 2);
    int64_t paths_total = 0;
    int pass = 0;
    const bool replay = graph != nullptr; // a captured graph stands in for the launches below
    if (!replay) {
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, ls, P.err, 1);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ls, P.counters, 8);
    }
    for (long long s0 = 0; !replay && s0 < p.spp; s0 += per_pass) {
        const long long k = std::min<long long>(per_pass, p.spp - s0);
        const long long n_paths = k * npix;
        RenderConsts rc;
        rc.n_paths = int32_t(n_paths);
        rc.npix = int32_t(npix);
        rc.width = camera->width;
        rc.seed = p.seed;
        rc.sample0 = int32_t(s0);
        rc.sample_begin = P.sample_begin;
        rc.max_bounces = p.max_bounces;
        rc.rr_min_bounces = p.rr_min_bounces;
        rc.light_alias = p.light_sampler == BB_LIGHTS_ALIAS ? 1 : 0;
        rc.train_stride = stride;
        paths_total += n_paths;
        hipEvent_t a = timed ? mark() : nullptr;
        hipLaunchKernelGGL(k_zero_ints, dim3((nslots + 255) / 256), dim3(256), 0, ls, P.counts, nslots);
        const int n_train = cache ? int((n_paths + stride - 1) / stride) : 0;
        if (cache) {
            hipLaunchKernelGGL(k_cache_pass_begin, dim3(1), dim3(64), 0, ls, Cv);
            hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, ls, Cv.counters, 4);
            hipLaunchKernelGGL(k_zero_ints, dim3((n_train + 255) / 256), dim3(256), 0, ls, Cv.tcount, n_train);
        }
        hipLaunchKernelGGL(k_raygen, dim3(unsigned((n_paths + kShadeBlock - 1) / kShadeBlock)), dim3(kShadeBlock), 0, ls,
                           cam, rc, P);
        if (timed) spans.push_back({ a, mark(), 2 });
        const int g_shade = grid_for(n_paths, kShadeBlock);
        for (int depth = 0; depth <= p.max_bounces; ++depth) {
            const int in = depth & 1;
            a = timed ? mark() : nullptr;
            const ClosestSrc csrc{ P.ray[in], P.hit };
            hipLaunchKernelGGL(k_closest, dim3(g_trace), dim3(tblock), lds, ls, V, P.counts + 4 * depth,
                               P.counts + 4 * depth + 2, csrc, P.gstack, lds_entries, refill_min, P.counters, P.err);
            hipEvent_t b = timed ? mark() : nullptr;
            if (timed) spans.push_back({ a, b, 0 });
            if (split && depth > 0) {
                HIP_OK(hipStreamWaitEvent(ls, x->shadowed, 0)); // the previous depth's shadow rays
                if (timed) b = mark();
            }
            hipLaunchKernelGGL(k_shade_fn, dim3(g_shade), dim3(kShadeBlock), 0, ls, V, rc, P, depth, Cv);
            hipEvent_t c2 = timed ? mark() : nullptr;
            if (timed) spans.push_back({ b, c2, 2 });
            const ShadowSrc ssrc{ P.sray, P.scontrib, P.L };
            if (split) {
                HIP_OK(hipEventRecord(x->shaded, ls));
                HIP_OK(hipStreamWaitEvent(x->side, x->shaded, 0));
                hipEvent_t s0 = timed ? mark_on(x->side) : nullptr;
                hipLaunchKernelGGL(k_shadow, dim3(g_shadow), dim3(tblock), lds_shadow, x->side, V,
                                   P.counts + 4 * depth + 1, P.counts + 4 * depth + 3, ssrc, P.gstack + ov_half,
                                   lds_entries, refill_shadow, P.counters, P.err);
                if (timed) spans.push_back({ s0, mark_on(x->side), 1 });
                HIP_OK(hipEventRecord(x->shadowed, x->side));
                continue;
            }
            hipLaunchKernelGGL(k_shadow, dim3(g_shadow), dim3(tblock), lds_shadow, ls, V, P.counts + 4 * depth + 1,
                               P.counts + 4 * depth + 3, ssrc, P.gstack, lds_entries, refill_shadow, P.counters, P.err);
            if (timed) spans.push_back({ c2, mark(), 1 });
        }
        if (split) HIP_OK(hipStreamWaitEvent(ls, x->shadowed, 0));
        if (cache) {
            // cache queries of the rendering paths, then this pass's training
            // rows; every count stays on the device
            // spans: the network's inference (3), the predictions applied and
            // the training rows built (5), the optimizer steps (4)
            hipEvent_t q0 = cache_timed ? mark() : nullptr;
            const int rq = nrc_query_counted(cache->nrc, Cv.qraw, Cv.counters, Cv.max_queries, Cv.qpred, st);
            if (rq != BBGPU_OK) return rq;
            hipEvent_t q1 = cache_timed ? mark() : nullptr;
            hipLaunchKernelGGL(k_cache_apply, dim3(std::min(grid_for(n_paths, kShadeBlock), 4096)), dim3(kShadeBlock), 0,
                               st, Cv, P.L, pass);
            const int batch = cache->cp.train_batch > 0 ? cache->cp.train_batch : 65536;
            const int used_rows = int(std::min<long long>(Cv.max_rows, (long long)cache->cp.train_steps * batch));
            hipLaunchKernelGGL(k_cache_targets, dim3((n_train + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0,
                               st, Cv, P.L, n_train, stride, used_rows);
            hipEvent_t q2 = cache_timed ? mark() : nullptr;
            for (int s1 = 0; s1 < cache->cp.train_steps; ++s1) {
                const int rt = nrc_train_counted(cache->nrc, Cv.rows, Cv.targets, Cv.counters + 1, Cv.max_rows, batch,
                                                 s1, st);
                if (rt != BBGPU_OK) return rt;
            }
            hipEvent_t q3 = cache_timed ? mark() : nullptr;
            hipLaunchKernelGGL(k_cache_tally, dim3(1), dim3(64), 0, st, Cv, cache->cp.train_steps, batch);
            if (cache_timed) {
                spans.push_back({ q0, q1, 3 });
                spans.push_back({ q1, q2, 5 });
                spans.push_back({ q2, q3, 4 });
                spans.push_back({ q3, mark(), 5 });
            }
        }
        a = timed ? mark() : nullptr;
        hipLaunchKernelGGL(k_film, dim3((npix + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, ls, rc, P.L,
                           film_sum_dev, film_count_dev, cache ? Cv.state + 1 : nullptr);
        ++pass;
        if (timed) spans.push_back({ a, mark(), 2 });
        if (!capturing) HIP_OK(hipGetLastError());
        if (timed) {
            // ray counts of this pass (counts are reused by the next pass)
            std::vector<int> h(nslots);
            HIP_OK(hipMemcpyAsync(h.data(), P.counts, sizeof(int) * nslots, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            for (int dd = 0; dd <= p.max_bounces; ++dd) {
                stats->rays_closest += h[4 * dd];
                stats->rays_shadow += h[4 * dd + 1];
            }
        }
    }
    if (capturing) {
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(x->capture, &g);
        hipGraphExec_t e = nullptr;
        if (ec != hipSuccess || !g || hipGetLastError() != hipSuccess ||
            hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) hipGraphDestroy(g);
            bb::set_error("bbgpu_render: graph capture failed");
            return BBGPU_ERR_DEVICE;
        }
        if (x->graphs.size() >= kMaxGraphs) {
            drop_graph(x->graphs.front());
            x->graphs.erase(x->graphs.begin());
        }
        x->graphs.push_back(GraphEntry{ std::move(key), g, e });
        graph = &x->graphs.back();
    }
    if (graph) HIP_OK(hipGraphLaunch(graph->exec, st));
    if (cache) {
        // one synchronization per call: the pass dropped for non-finite
        // predictions (rendered again without the cache, into the same film),
        // the divergence flag and the statistics
        int stv[4] = { 0, 0, 0, 0 };
        int flag = 0;
        unsigned long long tally[2] = { 0, 0 };
        float loss = 0.0f;
        HIP_OK(hipMemcpyAsync(stv, Cv.state, sizeof(stv), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(&flag, Cv.diverged, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(tally, Cv.tally, sizeof(tally), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(&loss, nrc_loss(cache->nrc), sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        nrc_note_diverged(cache->nrc, flag != 0);
        if (stv[2] > 0) {
            const long long s0 = (long long)(stv[2] - 1) * per_pass;
            BbRenderParams rp = p;
            rp.sample_begin = p.sample_begin + int32_t(s0);
            rp.spp = int32_t(std::min<long long>(per_pass, p.spp - s0));
            rp.graph = 0;
            const int rr = render_impl(scene, camera, &rp, film_sum_dev, film_count_dev, stream, nullptr, 0, nullptr);
            if (rr != BBGPU_OK) return rr;
        }
        if (cache->stats) {
            cache->stats->queries = int64_t(tally[0]);
            cache->stats->train_rows = int64_t(tally[1]);
            cache->stats->loss = loss;
            cache->stats->diverged = flag != 0 ? 1 : 0;
            cache->stats->passes_redone = stv[2] > 0 ? 1 : 0;
            for (const Span& sp : spans) {
                float d = 0.0f;
                hipEventElapsedTime(&d, sp.a, sp.b);
                if (sp.kind == 3) cache->stats->ms_query += d;
                if (sp.kind == 4) cache->stats->ms_train += d;
                if (sp.kind == 5) cache->stats->ms_rows += d;
            }
        }
    }
    hipEvent_t t_end = (timed || (cache && cache->stats)) ? mark() : nullptr;
    if (host_film) {
        HIP_OK(hipMemcpyAsync(film_sum_user, film_sum_dev, sum_bytes, hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(film_count_user, film_count_dev, count_bytes, hipMemcpyDeviceToHost, st));
    }
    if (!wait && !timed && !host_film && !cache) return BBGPU_OK;
    int err = 0;
    unsigned long long c[8] = { 0 };
    HIP_OK(hipMemcpyAsync(&err, P.err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(c, P.counters, sizeof(c), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (err) {
        bb::set_error("bbgpu_render: traversal stack overflow");
        return BBGPU_ERR_DEVICE;
    }
    if (timed) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, t_begin, t_end);
        stats->ms_total = ms;
        stats->ms_trace_closest = stats->ms_trace_shadow = stats->ms_shade = 0.0;
        stats->launches_closest = stats->launches_shadow = 0;
        for (const Span& sp : spans) {
            float d = 0.0f;
            hipEventElapsedTime(&d, sp.a, sp.b);
            if (sp.kind == 0) {
                stats->ms_trace_closest += d;
                stats->launches_closest++;
            } else if (sp.kind == 1) {
                stats->ms_trace_shadow += d;
                stats->launches_shadow++;
            } else if (sp.kind == 2) {
                stats->ms_shade += d;
            }
        }
        stats->paths = paths_total;
        if (count) {
            stats->closest.rays = uint64_t(stats->rays_closest);
            stats->closest.nodes = c[0];
            stats->closest.prims = c[1];
            stats->closest.accepts = c[2];
            stats->shadow.rays = uint64_t(stats->rays_shadow);
            stats->shadow.nodes = c[3];
            stats->shadow.prims = c[4];
            stats->shadow.accepts = c[5];
        }
    }
    if (cache && cache->stats) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, t_begin_all, t_end);
        cache->stats->ms_total = ms;
    }
    return BBGPU_OK;
}

// the error flag of a render queued with wait = false (after its stream is synchronized)
int check_render(BbScene* scene)
{
    SceneExtra* x = extra_for(scene);
    int err = 0;
    HIP_OK(hipMemcpy(&err, x->ws.P.err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        bb::set_error("bbgpu_render_batch: traversal stack overflow");
        return BBGPU_ERR_DEVICE;
    }
    return BBGPU_OK;
}

std::vector<hipStream_t>& stream_pool()
{
    static std::vector<hipStream_t> pool;
    return pool;
}

double job_cost(const BbRenderJob& j)
{
    return double(j.camera.width) * j.camera.height * std::max(0, j.params.spp) * (j.params.max_bounces + 1);
}

} // namespace

extern "C" {

int bbgpu_render_batch(BbRenderJob* jobs, int32_t n_jobs, int32_t streams, BbBatchStats* stats)
{
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) {
        bb::set_error("bbgpu_render_batch: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    for (int32_t i = 0; i < n_jobs; ++i) {
        const BbRenderJob& j = jobs[i];
        if (!j.scene || !j.film_sum_dev || !j.film_count_dev) {
            bb::set_error("bbgpu_render_batch: job %d has a null scene or film", i);
            return BBGPU_ERR_INVALID;
        }
        if (!device_accessible(j.film_sum_dev) || !device_accessible(j.film_count_dev)) {
            bb::set_error("bbgpu_render_batch: job %d: films must be device memory", i);
            return BBGPU_ERR_INVALID;
        }
    }
    // 16 by default: four streams per hardware queue measured 5-8 % faster
    // than 8 on 32 uneven scenes
    const int ns = std::max(1, std::min(streams > 0 ? streams : 16, 16));
    std::vector<hipStream_t>& pool = stream_pool();
    while (int(pool.size()) < ns) {
        hipStream_t st = nullptr;
        HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); // default priority
        pool.push_back(st);
    }
    // jobs of one scene share its workspace and run in order on one stream;
    // scenes go to the least loaded stream, most expensive first
    std::vector<BbScene*> scenes;
    std::vector<double> scene_cost;
    for (int32_t i = 0; i < n_jobs; ++i) {
        auto it = std::find(scenes.begin(), scenes.end(), jobs[i].scene);
        if (it == scenes.end()) {
            scenes.push_back(jobs[i].scene);
            scene_cost.push_back(job_cost(jobs[i]));
        } else {
            scene_cost[size_t(it - scenes.begin())] += job_cost(jobs[i]);
        }
    }
    std::vector<size_t> order(scenes.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return scene_cost[a] > scene_cost[b]; });
    std::vector<double> load(size_t(ns), 0.0);
    std::vector<int> scene_stream(scenes.size(), 0);
    for (size_t oi : order) {
        const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
        scene_stream[oi] = k;
        load[size_t(k)] += scene_cost[oi];
    }
    HIP_OK(hipDeviceSynchronize());
    EventPool evp;
    std::vector<hipEvent_t> start(static_cast<size_t>(ns)), done(static_cast<size_t>(n_jobs), nullptr);
    std::vector<int> job_stream(static_cast<size_t>(n_jobs), 0);
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < ns; ++k) {
        start[size_t(k)] = evp.get();
        HIP_OK(hipEventRecord(start[size_t(k)], pool[size_t(k)]));
    }
    int64_t samples = 0;
    // every scene keeps its own path state: the automatic pass size shares a
    // quarter of the device memory among the batch's scenes
    const long long auto_paths = std::max<long long>(
        1ll << 22, std::min<long long>(1ll << 28, device_memory() / 4 / 192 / (long long)scenes.size()));
    for (size_t oi : order)
        for (int32_t i = 0; i < n_jobs; ++i) {
            if (jobs[i].scene != scenes[oi]) continue;
            hipStream_t st = pool[size_t(scene_stream[oi])];
            BbRenderParams jp = jobs[i].params;
            if (jp.wavefront_paths <= 0) jp.wavefront_paths = int32_t(auto_paths);
            const int rc = render_impl(jobs[i].scene, &jobs[i].camera, &jp, jobs[i].film_sum_dev,
                                       jobs[i].film_count_dev, st, nullptr, 0, nullptr, false, false);
            if (rc != BBGPU_OK) return rc;
            done[size_t(i)] = evp.get();
            job_stream[size_t(i)] = scene_stream[oi];
            HIP_OK(hipEventRecord(done[size_t(i)], st));
            samples += int64_t(jobs[i].camera.width) * jobs[i].camera.height * jobs[i].params.spp;
        }
    for (int k = 0; k < ns; ++k) HIP_OK(hipStreamSynchronize(pool[size_t(k)]));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats && stats->job_ms)
        for (int32_t i = 0; i < n_jobs; ++i) {
            float e = 0.0f;
            HIP_OK(hipEventElapsedTime(&e, start[size_t(job_stream[size_t(i)])], done[size_t(i)]));
            stats->job_ms[i] = e;
        }
    for (BbScene* sc : scenes) {
        const int rc = check_render(sc);
        if (rc != BBGPU_OK) return rc;
    }
    if (stats) {
        stats->ms_total = ms;
        stats->samples = samples;
        stats->streams = ns;
    }
    return BBGPU_OK;
}

int bbgpu_autotune(BbScene* scene, const BbCameraDesc* camera, const BbRenderParams* params, int32_t probe_spp,
                   BbRenderParams* tuned_out)
{
    if (!scene || !camera || !params || !tuned_out) {
        bb::set_error("bbgpu_autotune: null argument");
        return BBGPU_ERR_INVALID;
    }
    // coordinate descent over the knobs that change speed, not results:
    // refill threshold, LDS stack entries, paths in flight
    BbRenderParams best = *params;
    best.spp = std::max(1, std::min(probe_spp > 0 ? probe_spp : 2, std::max(params->spp, 1)));
    const long long npix = (long long)camera->width * camera->height;
    float *sum = nullptr;
    int32_t* cnt = nullptr;
    HIP_OK(hipMalloc(&sum, sizeof(float) * 3 * size_t(npix)));
    if (hipMalloc(&cnt, sizeof(int32_t) * size_t(npix)) != hipSuccess) {
        (void)hipFree(sum);
        bb::set_error("bbgpu_autotune: film allocation failed");
        return BBGPU_ERR_NOMEM;
    }
    auto time_of = [&](const BbRenderParams& q, double& ms) {
        BbRenderStats st;
        const int rc = render_impl(scene, camera, &q, sum, cnt, nullptr, &st, 0, nullptr);
        ms = st.ms_total;
        return rc;
    };
    double best_ms = 0.0;
    int rc = time_of(best, best_ms); // warm-up, then the baseline
    if (rc == BBGPU_OK) rc = time_of(best, best_ms);
    // one knob at a time, keeping the best so far: refill threshold, block
    // size, stack placement (LDS entries, or the private scratch stack),
    // paths per pass
    const int refills[] = { 16, 32, 48 };
    const int blocks[] = { 64, 128, 256 };
    const int lds[] = { 8, 16, 24, -1 }; // -1: scratch stack
    const long long paths[] = { 1ll << 22, 1ll << 24, 1ll << 26 };
    const int n_values[4] = { 3, 3, 4, 3 };
    for (int knob = 0; knob < 4 && rc == BBGPU_OK; ++knob)
        for (int v = 0; v < n_values[knob] && rc == BBGPU_OK; ++v) {
            BbRenderParams q = best;
            if (knob == 0) q.trace_refill = q.shadow_refill = refills[v];
            if (knob == 1) q.trace_block = blocks[v];
            if (knob == 2) {
                q.trace_stack = lds[v] < 0 ? BB_STACK_SCRATCH : BB_STACK_LDS;
                q.trace_lds_stack = lds[v] < 0 ? 0 : lds[v];
            }
            if (knob == 3) q.wavefront_paths = int32_t(std::min<long long>(paths[v], 1ll << 30));
            double ms = 0.0;
            rc = time_of(q, ms);
            if (rc == BBGPU_ERR_UNSUPPORTED) { // a variant this scene cannot use
                rc = BBGPU_OK;
                continue;
            }
            if (rc == BBGPU_OK && ms < best_ms) {
                best_ms = ms;
                best = q;
            }
        }
    (void)hipFree(sum);
    (void)hipFree(cnt);
    if (rc != BBGPU_OK) return rc;
    *tuned_out = *params;
    tuned_out->trace_refill = best.trace_refill;
    tuned_out->shadow_refill = best.shadow_refill;
    tuned_out->trace_block = best.trace_block;
    tuned_out->trace_stack = best.trace_stack;
    tuned_out->trace_lds_stack = best.trace_lds_stack;
    tuned_out->wavefront_paths = best.wavefront_paths;
    return BBGPU_OK;
}

int bbgpu_render(BbScene* scene, const BbCameraDesc* camera, const BbRenderParams* params, float* film_sum_dev,
                 int32_t* film_count_dev, void* stream, BbRenderStats* stats, int32_t count_traversal)
{
    return render_impl(scene, camera, params, film_sum_dev, film_count_dev, stream, stats, count_traversal, nullptr);
}

int bbgpu_render_cached(BbScene* scene, BbNrc* nrc, const BbCameraDesc* camera, const BbRenderParams* params,
                        const BbCacheParams* cache_params, float* film_sum_dev, int32_t* film_count_dev, void* stream,
                        BbCacheStats* stats)
{
    if (!nrc || !cache_params) {
        bb::set_error("bbgpu_render_cached: null argument");
        return BBGPU_ERR_INVALID;
    }
    const BbCacheParams& c = *cache_params;
    if (c.query_bounce < 1 || c.train_stride < 1 || c.train_extra_bounces < 0 || c.train_batch < 0 || c.train_steps < 0 ||
        c.query_first < 0) {
        bb::set_error("bbgpu_render_cached: invalid cache parameters");
        return BBGPU_ERR_INVALID;
    }
    if (nrc_diverged(nrc)) {
        // the plain tracer is the fallback of a diverged cache
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->diverged = 1;
        }
        return render_impl(scene, camera, params, film_sum_dev, film_count_dev, stream, nullptr, 0, nullptr);
    }
    CacheCtx ctx{ nrc, c, stats };
    return render_impl(scene, camera, params, film_sum_dev, film_count_dev, stream, nullptr, 0, &ctx);
}

} // extern "C"

// releases the per-scene render workspace; called from bbgpu_scene_destroy
namespace bb
{
void release_scene_workspace(BbScene* s)
{
    std::lock_guard<std::mutex> g(g_extras_lock);
    for (size_t i = 0; i < g_extras.size(); ++i)
        if (g_extras[i].first == s) {
            for (GraphEntry& ge : g_extras[i].second->graphs) drop_graph(ge);
            if (g_extras[i].second->capture) hipStreamDestroy(g_extras[i].second->capture);
            if (g_extras[i].second->side) hipStreamDestroy(g_extras[i].second->side);
            if (g_extras[i].second->shaded) hipEventDestroy(g_extras[i].second->shaded);
            if (g_extras[i].second->shadowed) hipEventDestroy(g_extras[i].second->shadowed);
            free_workspace(g_extras[i].second->ws);
            if (g_extras[i].second->filter_cdf) hipFree(g_extras[i].second->filter_cdf);
            for (void* q : g_extras[i].second->cache.allocs) (void)hipFree(q);
            delete g_extras[i].second;
            g_extras.erase(g_extras.begin() + long(i));
            return;
        }
}
} // namespace bb

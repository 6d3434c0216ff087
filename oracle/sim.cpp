// The rigid-body World of phiray restated on the CPU: the checker the
// device solver (phiray_amd/csrc/world.hip) is held to.
//
// TEST INFRASTRUCTURE ONLY, like the rest of this folder: only the tests load
// it, and the product library never links or calls it.
//
// It restates phiray's XPBD world (src/sim/world.rs, contact.rs, broad.rs,
// shape.rs, math.rs, build.rs, narrow/closed.rs, narrow/sat.rs,
// narrow/manifold.rs) operation by operation: left-to-right sums, no fused
// multiply-add (built with -ffp-contract=off), ties in min/max selections as
// Rust's Iterator::min_by (first) and max_by (last) send them. The
// broadphase is brute force over the swept boxes (broad.rs walks a Bvh over
// the same boxes; the pair set does not depend on the tree).
//
// Shapes: box, sphere and capsule (closed forms, SAT with clipping). Hulls,
// cylinders, cones and triangle meshes, which phiray takes through GJK/EPA
// and its mesh midphase, are refused, as is a box-capsule pair.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_UNSUPPORTED (-3)
#define ORC_ERR_SIZE (-4)
#define ORC_CHECK(cond, code, msg) \
    do {                           \
        if (!(cond)) return (code); \
    } while (0)

namespace {

// ---- math (src/sim/math.rs) -------------------------------------------------
struct V3 {
    float v[3];
    float& operator[](int k) { return v[k]; }
    float operator[](int k) const { return v[k]; }
};
struct Q4 {
    float v[4];
    float& operator[](int k) { return v[k]; }
    float operator[](int k) const { return v[k]; }
};

inline V3 mk(float x, float y, float z) { return V3{{x, y, z}}; }
inline V3 add(V3 a, V3 b) { return mk(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 sub(V3 a, V3 b) { return mk(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 scale(V3 a, float s) { return mk(a[0] * s, a[1] * s, a[2] * s); }
inline V3 mul3(V3 a, V3 b) { return mk(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
inline float dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(V3 a, V3 b) {
    return mk(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline float len2(V3 a) { return dot(a, a); }
inline float len(V3 a) { return std::sqrt(len2(a)); }
inline V3 norm(V3 a) {
    float l = len(a);
    return l > 0.0f ? scale(a, 1.0f / l) : mk(0, 0, 0);
}
inline V3 neg(V3 a) { return mk(-a[0], -a[1], -a[2]); }
// Rust's f32::min / max: the other operand when one is NaN.
inline float rmin(float a, float b) { return std::fmin(a, b); }
inline float rmax(float a, float b) { return std::fmax(a, b); }
inline float rclamp(float x, float lo, float hi) { return rmin(rmax(x, lo), hi); }
inline float copysign_(float m, float s) { return std::copysign(m, s); }

inline Q4 q_mul(Q4 a, Q4 b) {
    return Q4{{a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1], a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
               a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3], a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]}};
}
inline Q4 q_conj(Q4 q) { return Q4{{-q[0], -q[1], -q[2], q[3]}}; }
inline Q4 q_norm(Q4 q) {
    float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (l > 0.0f) return Q4{{q[0] / l, q[1] / l, q[2] / l, q[3] / l}};
    return Q4{{0, 0, 0, 1}};
}
inline V3 rotate(Q4 q, V3 v) {  // q v q*
    V3 u = mk(q[0], q[1], q[2]);
    V3 t = scale(cross(u, v), 2.0f);
    return add(add(v, scale(t, q[3])), cross(u, t));
}
inline V3 unrotate(Q4 q, V3 v) { return rotate(q_conj(q), v); }
inline Q4 q_integrate(Q4 q, V3 w, float h) {
    Q4 d = q_mul(Q4{{w[0], w[1], w[2], 0.0f}}, q);
    return q_norm(Q4{{q[0] + 0.5f * h * d[0], q[1] + 0.5f * h * d[1], q[2] + 0.5f * h * d[2], q[3] + 0.5f * h * d[3]}});
}
inline V3 q_delta_w(Q4 q, Q4 prev, float h) {
    Q4 d = q_mul(q, q_conj(prev));
    V3 w = scale(mk(d[0], d[1], d[2]), 2.0f / h);
    return d[3] < 0.0f ? neg(w) : w;
}

// ---- state (src/sim/state.rs) -------------------------------------------------
enum : uint32_t { K_ACTIVE = 0, K_PASSIVE = 1, K_KINEMATIC = 2 };
enum : uint32_t { S_BOX = 0, S_SPHERE = 1, S_CAPSULE = 2, S_CYLINDER = 3, S_CONE = 4, S_HULL = 5, S_MESH = 6 };
constexpr uint32_t SLEEPING = 1u << 2, DEACTIVATION = 1u << 3, KIND_MASK = 3u;

struct Body {  // 64 B
    V3 x;
    float inv_mass;
    Q4 q;
    V3 v;
    uint32_t shape;
    V3 w;
    uint32_t flags;
    uint32_t kind() const { return (flags & KIND_MASK) == 1 ? K_PASSIVE : (flags & KIND_MASK) == 2 ? K_KINEMATIC : K_ACTIVE; }
    bool asleep() const { return (flags & SLEEPING) != 0; }
    void set_asleep(bool a) { flags = a ? (flags | SLEEPING) : (flags & ~SLEEPING); }
    bool dynamic() const { return kind() == K_ACTIVE && !asleep(); }
    bool moves() const { return kind() != K_PASSIVE && !asleep(); }
};
struct BodyConst {  // 48 B
    V3 inv_inertia;
    float friction;
    V3 com_local;
    float restitution, lin_damp, ang_damp, margin;
    uint32_t groups;
};
struct BodyPrev {
    V3 x;
    float _a;
    Q4 q;
};
struct Contact {  // 56 B
    uint32_t a, b;
    V3 anchor_a;
    float lambda_n;
    V3 anchor_b;
    float lambda_t;
    V3 normal;
    float depth0;
};
static_assert(sizeof(Body) == 64 && sizeof(BodyConst) == 48 && sizeof(Contact) == 56, "state layouts");

struct Aabb {
    V3 lo, hi;
};

// ---- shapes (src/sim/shape.rs) --------------------------------------------------
V3 support(const orc_shape& s, V3 d) {
    switch (s.kind) {
        case S_SPHERE: return scale(norm(d), s.radius);
        case S_BOX: return mk(copysign_(s.half[0], d[0]), copysign_(s.half[1], d[1]), copysign_(s.half[2], d[2]));
        default:  // capsule
            return add(scale(norm(d), s.radius), mk(0.0f, 0.0f, copysign_(s.half[2], d[2])));
    }
}
V3 support_world(const orc_shape& s, V3 x, Q4 q, V3 d) { return add(x, rotate(q, support(s, unrotate(q, d)))); }

Aabb shape_bounds(const orc_shape& s, V3 x, Q4 q) {
    auto axis = [&](int k, float sign) {
        V3 e = mk(0, 0, 0);
        e[k] = sign;
        return support_world(s, x, q, e)[k];
    };
    return Aabb{mk(axis(0, -1.0f), axis(1, -1.0f), axis(2, -1.0f)), mk(axis(0, 1.0f), axis(1, 1.0f), axis(2, 1.0f))};
}

V3 inertia(const orc_shape& s, float mass) {
    auto boxed = [&](V3 h) {
        float x = 4.0f * h[0] * h[0], y = 4.0f * h[1] * h[1], z = 4.0f * h[2] * h[2];
        return scale(mk(y + z, x + z, x + y), mass / 12.0f);
    };
    const float r2 = s.radius * s.radius;
    switch (s.kind) {
        case S_SPHERE: {
            float i = 0.4f * mass * r2;
            return mk(i, i, i);
        }
        case S_CAPSULE: return boxed(mk(s.radius, s.radius, s.half[2] + s.radius));
        default: return boxed(mk(s.half[0], s.half[1], s.half[2]));
    }
}

void segment(const orc_shape& s, V3 x, Q4 q, V3& p0, V3& p1) {
    V3 h = rotate(q, mk(0.0f, 0.0f, s.half[2]));
    p0 = sub(x, h);
    p1 = add(x, h);
}

// ---- manifold (src/sim/narrow/manifold.rs) ---------------------------------------
struct Point {
    V3 a, b;
};
struct Manifold {
    V3 normal = mk(0, 0, 0);
    Point points[4];
    int n = 0;
    void push(V3 a, V3 b) {
        if (n < 4) points[n++] = Point{a, b};
    }
    static Manifold one(V3 normal, V3 a, V3 b) {
        Manifold m;
        m.normal = normal;
        m.push(a, b);
        return m;
    }
    Manifold flip() const {
        Manifold m = *this;
        m.normal = neg(normal);
        for (int i = 0; i < m.n; ++i) std::swap(m.points[i].a, m.points[i].b);
        return m;
    }
    float depth(int i) const { return dot(sub(points[i].a, points[i].b), normal); }
};

// The index Rust's max_by picks: the last of the equal maxima (NaN scores compare equal).
template <class F>
int last_max(int count, F score) {
    int best = 0;
    for (int i = 1; i < count; ++i) {
        float a = score(best), b = score(i);
        if (!(b < a)) best = i;  // Ordering::Equal (incl. NaN) and Greater move to the later one
    }
    return best;
}
// The index Rust's min_by picks: the first of the equal minima.
template <class F>
int first_min(int count, F score) {
    int best = 0;
    for (int i = 1; i < count; ++i)
        if (score(i) < score(best)) best = i;
    return best;
}

Manifold gather(V3 normal, const std::vector<Point>& cands) {
    std::vector<Point> pts;
    for (const Point& c : cands) {
        bool dup = false;
        for (const Point& p : pts)
            if (len2(sub(p.a, c.a)) < 1e-8f) { dup = true; break; }
        if (!dup) pts.push_back(c);
    }
    Manifold m;
    m.normal = normal;
    if (pts.size() <= 4) {
        for (const Point& p : pts) m.push(p.a, p.b);
        return m;
    }
    const int np = (int)pts.size();
    auto depth = [&](const Point& p) { return dot(sub(p.a, p.b), normal); };
    const int i0 = last_max(np, [&](int i) { return depth(pts[i]); });
    const Point p0 = pts[i0];
    const int i1 = last_max(np, [&](int i) { return len2(sub(pts[i].a, p0.a)); });
    const Point p1 = pts[i1];
    const int i2 = last_max(np, [&](int i) { return len2(cross(sub(pts[i].a, p0.a), sub(pts[i].a, p1.a))); });
    const Point p2 = pts[i2];
    const int i3 = last_max(np, [&](int i) {
        float best = 3.40282347e38f;
        for (const Point* c : {&p0, &p1, &p2}) best = rmin(best, len2(sub(pts[i].a, c->a)));
        return best;
    });
    int chosen[4] = {i0, i1, i2, i3}, nc = 0, out[4];
    for (int k = 0; k < 4; ++k)  // Vec::dedup: consecutive repeats only
        if (nc == 0 || out[nc - 1] != chosen[k]) out[nc++] = chosen[k];
    for (int k = 0; k < nc; ++k) m.push(pts[out[k]].a, pts[out[k]].b);
    return m;
}

// ---- closed forms (src/sim/narrow/closed.rs) -----------------------------------------
Manifold spheres(V3 ca, float ra, V3 cb, float rb, float slop) {
    V3 d = sub(cb, ca);
    float dist = len(d);
    if (dist - ra - rb > slop) return Manifold();
    V3 n = dist > 0.0f ? scale(d, 1.0f / dist) : mk(0, 0, 1);
    return Manifold::one(n, add(ca, scale(n, ra)), sub(cb, scale(n, rb)));
}

V3 closest_on_segment(V3 p, V3 a, V3 b) {
    V3 ab = sub(b, a);
    float l2 = len2(ab);
    float t = l2 > 0.0f ? rclamp(dot(sub(p, a), ab) / l2, 0.0f, 1.0f) : 0.0f;
    return add(a, scale(ab, t));
}

void closest_segments(V3 p1, V3 q1, V3 p2, V3 q2, V3& o1, V3& o2) {
    V3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    float a = len2(d1), e = len2(d2), f = dot(d2, r);
    const float eps = 1e-12f;
    float s, t;
    if (a <= eps && e <= eps) {
        o1 = p1;
        o2 = p2;
        return;
    }
    if (a <= eps) {
        s = 0.0f;
        t = rclamp(f / e, 0.0f, 1.0f);
    } else {
        float c = dot(d1, r);
        if (e <= eps) {
            s = rclamp(-c / a, 0.0f, 1.0f);
            t = 0.0f;
        } else {
            float b = dot(d1, d2);
            float denom = a * e - b * b;
            s = denom != 0.0f ? rclamp((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                s = rclamp(-c / a, 0.0f, 1.0f);
                t = 0.0f;
            }
            if (t > 1.0f) {
                s = rclamp((b - c) / a, 0.0f, 1.0f);
                t = 1.0f;
            }
        }
    }
    o1 = add(p1, scale(d1, s));
    o2 = add(p2, scale(d2, t));
}

Manifold sphere_capsule(V3 c, float r, V3 s0, V3 s1, float rc, float slop) {
    return spheres(c, r, closest_on_segment(c, s0, s1), rc, slop);
}

Manifold capsules(V3 a0, V3 a1, float ra, V3 b0, V3 b1, float rb, float slop) {
    V3 pa, pb;
    closest_segments(a0, a1, b0, b1, pa, pb);
    return spheres(pa, ra, pb, rb, slop);
}

Manifold sphere_box(V3 c, float r, V3 x, Q4 q, V3 half, float slop) {
    V3 cl = unrotate(q, sub(c, x));
    V3 p = mk(rclamp(cl[0], -half[0], half[0]), rclamp(cl[1], -half[1], half[1]), rclamp(cl[2], -half[2], half[2]));
    V3 d = sub(cl, p);
    float dist = len(d);
    if (dist == 0.0f) {  // centre inside: out through the nearest face
        V3 gap = mk(half[0] - std::fabs(cl[0]), half[1] - std::fabs(cl[1]), half[2] - std::fabs(cl[2]));
        int k = first_min(3, [&](int i) { return gap[i]; });
        V3 out = mk(0, 0, 0);
        out[k] = copysign_(1.0f, cl[k]);
        V3 face = cl;
        face[k] = copysign_(half[k], cl[k]);
        V3 n = neg(rotate(q, out));
        return Manifold::one(n, add(c, scale(n, r)), add(x, rotate(q, face)));
    }
    if (dist - r > slop) return Manifold();
    V3 n = neg(rotate(q, scale(d, 1.0f / dist)));
    return Manifold::one(n, add(c, scale(n, r)), add(x, rotate(q, p)));
}

// ---- SAT with clipping (src/sim/narrow/sat.rs), boxes only --------------------------
struct Face {
    V3 n;
    uint8_t v[4];
    int nv;
};
struct Poly {
    V3 v[8];
    Face faces[6];
    uint8_t edges[12][2];
    int nv = 8, nf = 6, ne = 12, axes = 3;

    static Poly boxed(V3 x, Q4 q, V3 half) {
        Poly p;
        for (int i = 0; i < 8; ++i) {
            auto s = [&](int bit) { return ((i >> bit) & 1) == 1 ? 1.0f : -1.0f; };
            p.v[i] = add(x, rotate(q, mk(half[0] * s(0), half[1] * s(1), half[2] * s(2))));
        }
        auto axis = [&](int k, float sign) {
            V3 e = mk(0, 0, 0);
            e[k] = sign;
            return rotate(q, e);
        };
        const uint8_t fv[6][4] = {{1, 3, 7, 5}, {0, 4, 6, 2}, {2, 6, 7, 3}, {0, 1, 5, 4}, {4, 5, 7, 6}, {0, 2, 3, 1}};
        const int fa[6] = {0, 0, 1, 1, 2, 2};
        const float fs[6] = {1, -1, 1, -1, 1, -1};
        for (int f = 0; f < 6; ++f) {
            p.faces[f].n = axis(fa[f], fs[f]);
            std::memcpy(p.faces[f].v, fv[f], 4);
            p.faces[f].nv = 4;
        }
        const uint8_t e[12][2] = {{0, 1}, {0, 2}, {0, 4}, {2, 3}, {4, 5}, {6, 7}, {1, 3}, {4, 6}, {5, 7}, {1, 5}, {2, 6}, {3, 7}};
        std::memcpy(p.edges, e, sizeof e);
        return p;
    }
    void extent(V3 axis, float& lo, float& hi) const {
        lo = 3.40282347e38f;
        hi = -3.40282347e38f;
        for (int i = 0; i < nv; ++i) {
            float t = dot(v[i], axis);
            lo = rmin(lo, t);
            hi = rmax(hi, t);
        }
    }
    const Face& facing(V3 d) const {
        return faces[last_max(nf, [&](int i) { return dot(faces[i].n, d); })];
    }
    void edge(int e, V3& p, V3& q) const {
        p = v[edges[e][0]];
        q = v[edges[e][1]];
    }
    V3 edge_dir(int e) const { return sub(v[edges[e][1]], v[edges[e][0]]); }
    V3 face_centre(const Face& f) const {
        V3 a = mk(0, 0, 0);
        for (int k = 0; k < f.nv; ++k) a = add(a, v[f.v[k]]);
        return scale(a, 1.0f / (float)f.nv);
    }
};

void overlap(const Poly& a, const Poly& b, V3 u, float& o, V3& n) {
    float alo, ahi, blo, bhi;
    a.extent(u, alo, ahi);
    b.extent(u, blo, bhi);
    float o1 = ahi - blo, o2 = bhi - alo;
    if (o1 <= o2) o = o1, n = u;
    else o = o2, n = neg(u);
}

std::vector<V3> clip_plane(const std::vector<V3>& poly, V3 p0, V3 s) {
    std::vector<V3> out;
    const size_t np = poly.size();
    for (size_t i = 0; i < np; ++i) {
        V3 cur = poly[i], next = poly[(i + 1) % np];
        float dc = dot(sub(cur, p0), s), dn = dot(sub(next, p0), s);
        if (dc >= -1e-5f) out.push_back(cur);
        if ((dc >= -1e-5f) != (dn >= -1e-5f)) out.push_back(add(cur, scale(sub(next, cur), dc / (dc - dn))));
    }
    return out;
}

Manifold clip(const Poly& a, const Poly& b, V3 n, float slop, bool on_b) {
    const Poly& r = on_b ? b : a;
    const Poly& in = on_b ? a : b;
    const V3 m = on_b ? neg(n) : n;
    const Face& rf = r.facing(m);
    const Face& inc = in.facing(neg(m));
    std::vector<V3> poly;
    for (int k = 0; k < inc.nv; ++k) poly.push_back(in.v[inc.v[k]]);
    V3 rc = r.face_centre(rf);
    for (int e = 0; e < rf.nv; ++e) {
        V3 p0 = r.v[rf.v[e]], p1 = r.v[rf.v[(e + 1) % rf.nv]];
        V3 s = norm(cross(rf.n, sub(p1, p0)));
        if (dot(s, sub(rc, p0)) < 0.0f) s = neg(s);
        poly = clip_plane(poly, p0, s);
    }
    V3 on = r.v[rf.v[0]];
    std::vector<Point> cands;
    for (V3 p : poly) {
        float depth = dot(sub(on, p), rf.n);
        if (!(depth > -slop)) continue;
        cands.push_back(on_b ? Point{p, sub(p, scale(n, depth))} : Point{add(p, scale(n, depth)), p});
    }
    return gather(n, cands);
}

Manifold edge_contact(const Poly& a, const Poly& b, int ea, int eb, V3 n) {
    auto pick = [&](const Poly& p, V3 dir, float sign, V3& e0, V3& e1) {
        auto along = [&](int e) {
            V3 p0, p1;
            p.edge(e, p0, p1);
            return sign * dot(add(p0, p1), n);
        };
        auto parallel = [&](int e) {
            V3 ed = p.edge_dir(e);
            return len2(cross(ed, dir)) < 1e-8f * len2(ed) * len2(dir);
        };
        int best = -1;
        for (int e = 0; e < p.ne; ++e) {
            if (!parallel(e)) continue;
            if (best < 0 || !(along(e) < along(best))) best = e;  // max_by: the last of equal maxima
        }
        p.edge(best < 0 ? 0 : best, e0, e1);
    };
    V3 a0, a1, b0, b1, pa, pb;
    pick(a, a.edge_dir(ea), 1.0f, a0, a1);
    pick(b, b.edge_dir(eb), -1.0f, b0, b1);
    closest_segments(a0, a1, b0, b1, pa, pb);
    return Manifold::one(n, pa, pb);
}

Manifold sat(const Poly& a, const Poly& b, float slop) {
    enum { FACE_A, FACE_B, EDGE };
    float best_o = 3.40282347e38f;
    V3 best_axis = mk(0, 0, 0);
    int src = FACE_A, sea = 0, seb = 0;
    auto consider = [&](V3 u, int source, int ea, int eb, float bias) {
        float o;
        V3 n;
        overlap(a, b, u, o, n);
        if (o + bias < best_o) best_o = o, best_axis = n, src = source, sea = ea, seb = eb;
        return o >= -slop;
    };
    for (int f = 0; f < a.nf; ++f)
        if (!consider(a.faces[f].n, FACE_A, 0, 0, 0.0f)) return Manifold();
    for (int f = 0; f < b.nf; ++f)
        if (!consider(b.faces[f].n, FACE_B, 0, 0, 0.0f)) return Manifold();
    for (int ea = 0; ea < a.axes; ++ea)
        for (int eb = 0; eb < b.axes; ++eb) {
            V3 u = cross(a.edge_dir(ea), b.edge_dir(eb));
            float l = len(u);
            if (l < 1e-6f) continue;
            if (!consider(scale(u, 1.0f / l), EDGE, ea, eb, 1e-3f)) return Manifold();
        }
    if (src == FACE_A) return clip(a, b, best_axis, slop, false);
    if (src == FACE_B) return clip(a, b, best_axis, slop, true);
    return edge_contact(a, b, sea, seb, best_axis);
}

}  // namespace

namespace {

// ---- narrowphase dispatch (src/sim/narrow/mod.rs) -----------------------------------
constexpr float SLOP = 0.01f;

struct View {
    const orc_shape* s;
    V3 x;
    Q4 q;
    float margin;
};

Manifold convex_pair(const View& a, const View& b, float extra) {
    const float slop = a.margin + b.margin + SLOP + extra;
    const uint32_t ka = a.s->kind, kb = b.s->kind;
    V3 a0, a1, b0, b1;
    if (ka == S_SPHERE && kb == S_SPHERE) return spheres(a.x, a.s->radius, b.x, b.s->radius, slop);
    if (ka == S_SPHERE && kb == S_CAPSULE) {
        segment(*b.s, b.x, b.q, b0, b1);
        return sphere_capsule(a.x, a.s->radius, b0, b1, b.s->radius, slop);
    }
    if (ka == S_CAPSULE && kb == S_SPHERE) return convex_pair(b, a, extra).flip();
    if (ka == S_CAPSULE && kb == S_CAPSULE) {
        segment(*a.s, a.x, a.q, a0, a1);
        segment(*b.s, b.x, b.q, b0, b1);
        return capsules(a0, a1, a.s->radius, b0, b1, b.s->radius, slop);
    }
    if (ka == S_SPHERE && kb == S_BOX) return sphere_box(a.x, a.s->radius, b.x, b.q, mk(b.s->half[0], b.s->half[1], b.s->half[2]), slop);
    if (ka == S_BOX && kb == S_SPHERE) return convex_pair(b, a, extra).flip();
    if (ka == S_BOX && kb == S_BOX)
        return sat(Poly::boxed(a.x, a.q, mk(a.s->half[0], a.s->half[1], a.s->half[2])),
                   Poly::boxed(b.x, b.q, mk(b.s->half[0], b.s->half[1], b.s->half[2])), slop);
    return Manifold();  // box-capsule goes through GJK/EPA in phiray: refused when the shape is added
}

// ---- contacts (src/sim/contact.rs) -------------------------------------------------
float inv_mass_along(const Body& b, const BodyConst& c, V3 r, V3 n) {
    if (b.inv_mass == 0.0f) return 0.0f;
    V3 rn = unrotate(b.q, cross(r, n));
    return b.inv_mass + dot(mul3(rn, c.inv_inertia), rn);
}
void arms(const Body& a, const Body& b, const Contact& c, V3& ra, V3& rb) {
    ra = rotate(a.q, c.anchor_a);
    rb = rotate(b.q, c.anchor_b);
}
V3 relative_velocity(const Body& a, const Body& b, V3 ra, V3 rb) { return sub(add(a.v, cross(a.w, ra)), add(b.v, cross(b.w, rb))); }

struct Share {
    V3 p, ra, rb, dir;
    float need;
};
struct Motion {
    V3 da, wa, db, wb;
};

Motion motion(const Body& a, const BodyConst& ka, const Body& b, const BodyConst& kb, const std::vector<Share>& shares) {
    auto turn = [](const Body& bd, const BodyConst& k, V3 r, V3 p) {
        return bd.inv_mass == 0.0f ? mk(0, 0, 0) : rotate(bd.q, mul3(unrotate(bd.q, cross(r, p)), k.inv_inertia));
    };
    Motion m{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
    for (const Share& s : shares) {
        m.da = add(m.da, scale(s.p, a.inv_mass));
        m.wa = add(m.wa, turn(a, ka, s.ra, s.p));
        m.db = sub(m.db, scale(s.p, b.inv_mass));
        m.wb = sub(m.wb, turn(b, kb, s.rb, s.p));
    }
    return m;
}

float fit(const Motion& m, const std::vector<Share>& shares) {
    float alpha = 1.0f;
    for (const Share& s : shares) {
        float got = dot(sub(add(m.da, cross(m.wa, s.ra)), add(m.db, cross(m.wb, s.rb))), s.dir);
        if (got * s.need > 0.0f) alpha = rmin(alpha, s.need / got);
    }
    return alpha;
}

float move_poses(Body& a, const BodyConst& ka, Body& b, const BodyConst& kb, const std::vector<Share>& shares) {
    Motion m = motion(a, ka, b, kb, shares);
    float alpha = fit(m, shares);
    a.x = add(a.x, scale(m.da, alpha));
    a.q = q_integrate(a.q, scale(m.wa, alpha), 1.0f);
    b.x = add(b.x, scale(m.db, alpha));
    b.q = q_integrate(b.q, scale(m.wb, alpha), 1.0f);
    return alpha;
}

void move_velocities(Body& a, const BodyConst& ka, Body& b, const BodyConst& kb, const std::vector<Share>& shares) {
    Motion m = motion(a, ka, b, kb, shares);
    float alpha = fit(m, shares);
    a.v = add(a.v, scale(m.da, alpha));
    a.w = add(a.w, scale(m.wa, alpha));
    b.v = add(b.v, scale(m.db, alpha));
    b.w = add(b.w, scale(m.wb, alpha));
}

// Runs of consecutive contacts of one pair.
template <class F>
void groups(std::vector<Contact>& cs, F each) {
    size_t at = 0;
    while (at < cs.size()) {
        size_t e = at;
        while (e < cs.size() && cs[e].a == cs[at].a && cs[e].b == cs[at].b) ++e;
        each(at, e);
        at = e;
    }
}

void presolve(const std::vector<Body>& bodies, std::vector<Contact>& cs, std::vector<float>& vn0) {
    for (size_t k = 0; k < cs.size(); ++k) {
        Contact& c = cs[k];
        const Body &a = bodies[c.a], &b = bodies[c.b];
        V3 ra, rb;
        arms(a, b, c, ra, rb);
        vn0[k] = dot(relative_velocity(a, b, ra, rb), c.normal);
        c.lambda_n = 0.0f;
        c.lambda_t = 0.0f;
    }
}

void project(std::vector<Body>& bodies, const std::vector<BodyConst>& consts, const std::vector<BodyPrev>& prev,
             std::vector<Contact>& cs, float cap, std::vector<V3>& push, std::vector<V3>& spin) {
    groups(cs, [&](size_t g0, size_t g1) {
        const uint32_t i = cs[g0].a, j = cs[g0].b;
        const BodyConst &ka = consts[i], &kb = consts[j];
        Body &a = bodies[i], &b = bodies[j];
        const size_t ng = g1 - g0;
        std::vector<Share> shares;
        std::vector<float> dls(ng, 0.0f), recovery(ng, 0.0f);
        std::vector<size_t> owners;
        for (size_t k = 0; k < ng; ++k) {
            const Contact& c = cs[g0 + k];
            V3 ra, rb;
            arms(a, b, c, ra, rb);
            V3 n = c.normal;
            float d = dot(sub(add(a.x, ra), add(b.x, rb)), n);
            float w = inv_mass_along(a, ka, ra, n) + inv_mass_along(b, kb, rb, n);
            if (d <= 0.0f || w == 0.0f) continue;
            V3 qa = add(prev[i].x, rotate(prev[i].q, c.anchor_a)), qb = add(prev[j].x, rotate(prev[j].q, c.anchor_b));
            float closed = rmax(dot(sub(sub(add(a.x, ra), qa), sub(add(b.x, rb), qb)), n), 0.0f);
            float hit = rmin(d, closed);
            float fix = hit + rmin(d - hit, cap);
            dls[k] = -fix / w;
            recovery[k] = fix > 0.0f ? (fix - hit) / fix : 0.0f;
            owners.push_back(k);
            shares.push_back(Share{scale(n, dls[k]), ra, rb, neg(n), fix});
        }
        if (!shares.empty()) {
            float alpha = move_poses(a, ka, b, kb, shares);
            for (size_t k = 0; k < ng; ++k) cs[g0 + k].lambda_n += dls[k] * alpha;
            std::vector<Share> only;
            bool any = false;
            for (size_t s = 0; s < shares.size(); ++s) {
                Share o = shares[s];
                o.p = scale(shares[s].p, recovery[owners[s]]);
                any = any || len2(o.p) > 0.0f;
                only.push_back(o);
            }
            if (any) {
                Motion m = motion(a, ka, b, kb, only);
                push[i] = add(push[i], scale(m.da, alpha));
                spin[i] = add(spin[i], scale(m.wa, alpha));
                push[j] = add(push[j], scale(m.db, alpha));
                spin[j] = add(spin[j], scale(m.wb, alpha));
            }
        }
        // Static friction, scaled by the manifold's Jacobi share before the cone.
        shares.clear();
        owners.clear();
        std::vector<V3> tangent(ng, mk(0, 0, 0));
        const float mu = ka.friction * kb.friction;
        for (size_t k = 0; k < ng; ++k) {
            dls[k] = 0.0f;
            const Contact& c = cs[g0 + k];
            if (c.lambda_n == 0.0f) continue;
            V3 ra, rb;
            arms(a, b, c, ra, rb);
            V3 pa = add(a.x, ra), pb = add(b.x, rb);
            V3 qa = add(prev[i].x, rotate(prev[i].q, c.anchor_a)), qb = add(prev[j].x, rotate(prev[j].q, c.anchor_b));
            V3 dp = sub(sub(pa, qa), sub(pb, qb));
            V3 dpt = sub(dp, scale(c.normal, dot(dp, c.normal)));
            float lt = len(dpt);
            if (lt < 1e-9f) continue;
            V3 t = scale(dpt, 1.0f / lt);
            float wt = inv_mass_along(a, ka, ra, t) + inv_mass_along(b, kb, rb, t);
            if (wt == 0.0f) continue;
            tangent[k] = t;
            dls[k] = -lt / wt;
            owners.push_back(k);
            shares.push_back(Share{scale(t, dls[k]), ra, rb, neg(t), lt});
        }
        if (shares.empty()) return;
        const float share = fit(motion(a, ka, b, kb, shares), shares);
        for (size_t s = 0; s < shares.size(); ++s) {
            const size_t k = owners[s];
            float want = dls[k] * share;
            float room = rmax(mu * std::fabs(cs[g0 + k].lambda_n) - std::fabs(cs[g0 + k].lambda_t), 0.0f);
            float sg = want > 0.0f ? 1.0f : (want < 0.0f ? -1.0f : (std::signbit(want) ? -1.0f : 1.0f));  // f32::signum
            if (want != want) sg = want;
            dls[k] = sg * rmin(std::fabs(want), room);
            shares[s].p = scale(tangent[k], dls[k]);
        }
        float alpha = move_poses(a, ka, b, kb, shares);
        for (size_t k = 0; k < ng; ++k) cs[g0 + k].lambda_t += dls[k] * alpha;
    });
}

void velocity_pass(std::vector<Body>& bodies, const std::vector<BodyConst>& consts, std::vector<Contact>& cs,
                   const std::vector<float>& vn0, float h, float g) {
    groups(cs, [&](size_t g0, size_t g1) {
        const uint32_t i = cs[g0].a, j = cs[g0].b;
        const BodyConst &ka = consts[i], &kb = consts[j];
        Body &a = bodies[i], &b = bodies[j];
        const float mu = ka.friction * kb.friction;
        std::vector<Share> shares;
        for (size_t k = g0; k < g1; ++k) {
            const Contact& c = cs[k];
            if (c.lambda_n == 0.0f) continue;
            V3 ra, rb;
            arms(a, b, c, ra, rb);
            V3 vrel = relative_velocity(a, b, ra, rb);
            V3 vt = sub(vrel, scale(c.normal, dot(vrel, c.normal)));
            float lt = len(vt);
            if (lt < 1e-9f) continue;
            V3 t = scale(vt, 1.0f / lt);
            float wt = inv_mass_along(a, ka, ra, t) + inv_mass_along(b, kb, rb, t);
            float jj = rmin(lt / wt, mu * std::fabs(c.lambda_n) / h);
            shares.push_back(Share{scale(t, -jj), ra, rb, neg(t), jj * wt});
        }
        if (!shares.empty()) move_velocities(a, ka, b, kb, shares);
        shares.clear();
        for (size_t k = g0; k < g1; ++k) {
            const Contact& c = cs[k];
            if (c.lambda_n == 0.0f) continue;
            V3 ra, rb;
            arms(a, b, c, ra, rb);
            V3 n = c.normal;
            float vn = dot(relative_velocity(a, b, ra, rb), n);
            float e = std::fabs(vn0[k]) <= 2.0f * g * h ? 0.0f : ka.restitution * kb.restitution;
            float wn = inv_mass_along(a, ka, ra, n) + inv_mass_along(b, kb, rb, n);
            float need = -e * rmax(vn0[k], 0.0f) - vn;
            if (wn > 0.0f && need < -1e-12f) shares.push_back(Share{scale(n, need / wn), ra, rb, n, need});
        }
        if (!shares.empty()) move_velocities(a, ka, b, kb, shares);
    });
}

// ---- broadphase (src/sim/broad.rs:52-91) -----------------------------------------------
// Every pair of swept boxes that overlap, by brute force: broad.rs walks a
// Bvh::build tree over the same boxes, and the pair set does not depend on the
// tree. Listed as (i, j), i < j, sorted, when either body moves and one is
// active and their collision groups meet.
constexpr float PAD = SLOP + 0.005f;

}  // namespace

// ---- the world (src/sim/world.rs, build.rs) --------------------------------------------
struct orc_world {
    std::vector<Body> bodies;
    std::vector<BodyConst> consts;
    std::vector<BodyPrev> prev;
    std::vector<orc_shape> shapes;
    std::vector<Contact> contacts;
    std::vector<float> vn0;
    std::vector<V3> push, spin;
    std::vector<uint32_t> objs;
    std::vector<V3> scales;
    std::vector<float> sleep_lin, sleep_ang, timers;
    V3 gravity = mk(0.0f, 0.0f, -9.81f);
    uint32_t substeps = 10, iterations = 10;
    float time_scale = 1.0f, max_recover = 2.0f, sleep_after = 2.0f;
    std::vector<std::pair<uint32_t, uint32_t>> last_pairs;
};

namespace {

void world_pairs(orc_world* w, const std::vector<Aabb>& boxes, std::vector<std::pair<uint32_t, uint32_t>>& out) {
    const uint32_t n = (uint32_t)w->bodies.size();
    out.clear();
    auto meets = [&](const Aabb& p, const Aabb& q) {
        return p.lo[0] <= q.hi[0] && p.hi[0] >= q.lo[0] && p.lo[1] <= q.hi[1] && p.hi[1] >= q.lo[1] && p.lo[2] <= q.hi[2] &&
               p.hi[2] >= q.lo[2];
    };
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j) {
            const Body &bi = w->bodies[i], &bj = w->bodies[j];
            if (!bi.moves() && !bj.moves()) continue;
            if (bi.kind() != K_ACTIVE && bj.kind() != K_ACTIVE) continue;
            if ((w->consts[i].groups & w->consts[j].groups) == 0u) continue;
            if (meets(boxes[i], boxes[j])) out.push_back({i, j});
        }
}

// Bounds, broadphase, narrowphase; a contact wakes what it touches (world.rs:131-151).
void collide(orc_world* w, float dt) {
    const size_t n = w->bodies.size();
    std::vector<Aabb> boxes(n);
    for (size_t i = 0; i < n; ++i) {  // broad.rs:52-70: swept by the motion over dt, grown by the margin
        const Body& b = w->bodies[i];
        Aabb now = shape_bounds(w->shapes[b.shape], b.x, b.q);
        V3 shift = scale(b.v, dt);
        V3 slo = add(now.lo, shift), shi = add(now.hi, shift);
        Aabb swept{mk(rmin(now.lo[0], slo[0]), rmin(now.lo[1], slo[1]), rmin(now.lo[2], slo[2])),
                   mk(rmax(now.hi[0], shi[0]), rmax(now.hi[1], shi[1]), rmax(now.hi[2], shi[2]))};
        float pad = w->consts[i].margin + PAD;
        boxes[i] = Aabb{mk(swept.lo[0] - pad, swept.lo[1] - pad, swept.lo[2] - pad), mk(swept.hi[0] + pad, swept.hi[1] + pad, swept.hi[2] + pad)};
    }
    std::vector<std::pair<uint32_t, uint32_t>>& pairs = w->last_pairs;
    world_pairs(w, boxes, pairs);
    w->contacts.clear();
    for (auto& pr : pairs) {
        const uint32_t i = pr.first, j = pr.second;
        const Body &bi = w->bodies[i], &bj = w->bodies[j];
        View a{&w->shapes[bi.shape], bi.x, bi.q, w->consts[i].margin}, b{&w->shapes[bj.shape], bj.x, bj.q, w->consts[j].margin};
        const float reach = rmax(len(sub(bi.v, bj.v)) * dt, 0.0f);
        Manifold m = convex_pair(a, b, reach);
        for (int k = 0; k < m.n; ++k) {
            const Point& p = m.points[k];
            Contact c{};
            c.a = i;
            c.b = j;
            c.anchor_a = unrotate(bi.q, sub(p.a, bi.x));
            c.anchor_b = unrotate(bj.q, sub(p.b, bj.x));
            c.normal = m.normal;
            c.depth0 = m.depth(k);
            w->contacts.push_back(c);
        }
    }
    for (const Contact& c : w->contacts) {
        const uint32_t ends[2][2] = {{c.a, c.b}, {c.b, c.a}};
        for (auto& e : ends)
            if (w->bodies[e[0]].asleep() && w->bodies[e[1]].moves()) {
                w->bodies[e[0]].set_asleep(false);
                w->timers[e[0]] = 0.0f;
            }
    }
    w->vn0.assign(w->contacts.size(), 0.0f);
    w->push.assign(n, mk(0, 0, 0));
    w->spin.assign(n, mk(0, 0, 0));
}

}  // namespace

extern "C" {

orc_world* orc_world_new(void) { return new orc_world(); }

void orc_world_free(orc_world* w) { delete w; }

// World::substeps / iterations / time_scale / max_recover / sleep_after / gravity.
void orc_world_configure(orc_world* w, uint32_t substeps, uint32_t iterations, float time_scale, float max_recover,
                         float sleep_after, const float gravity[3]) {
    w->substeps = substeps;
    w->iterations = iterations;
    w->time_scale = time_scale;
    w->max_recover = max_recover;
    w->sleep_after = sleep_after;
    if (gravity) w->gravity = mk(gravity[0], gravity[1], gravity[2]);
}

// World::add_shape (build.rs): box, sphere and capsule only.
int orc_world_add_shape(orc_world* w, const orc_shape* s, uint32_t* id) {
    ORC_CHECK(w && s && id, ORC_ERR_ARG, "orc_world_add_shape: null");
    ORC_CHECK(s->kind == S_BOX || s->kind == S_SPHERE || s->kind == S_CAPSULE, ORC_ERR_UNSUPPORTED,
              "orc_world_add_shape: hulls, cylinders, cones and meshes take GJK/EPA in phiray and are not restated");
    w->shapes.push_back(*s);
    *id = (uint32_t)w->shapes.size() - 1;
    return ORC_OK;
}

// World::add_body (build.rs:99-131).
int orc_world_add_body(orc_world* w, const orc_body_def* d, uint32_t* id) {
    ORC_CHECK(w && d && id, ORC_ERR_ARG, "orc_world_add_body: null");
    ORC_CHECK(d->shape < w->shapes.size(), ORC_ERR_ARG, "orc_world_add_body: no such shape");
    if (w->shapes[d->shape].kind == S_CAPSULE)
        for (const Body& b : w->bodies)
            ORC_CHECK(w->shapes[b.shape].kind != S_BOX, ORC_ERR_UNSUPPORTED, "orc_world_add_body: box-capsule pairs take GJK/EPA");
    if (w->shapes[d->shape].kind == S_BOX)
        for (const Body& b : w->bodies)
            ORC_CHECK(w->shapes[b.shape].kind != S_CAPSULE, ORC_ERR_UNSUPPORTED, "orc_world_add_body: box-capsule pairs take GJK/EPA");
    const bool movable = d->kind == K_ACTIVE && d->mass > 0.0f;
    const float inv_mass = movable ? 1.0f / d->mass : 0.0f;
    V3 in = inertia(w->shapes[d->shape], d->mass);
    V3 inv_in = mk(movable && in[0] > 0.0f ? 1.0f / in[0] : 0.0f, movable && in[1] > 0.0f ? 1.0f / in[1] : 0.0f,
                   movable && in[2] > 0.0f ? 1.0f / in[2] : 0.0f);
    uint32_t flags = d->kind;
    if (d->has_sleep) flags |= DEACTIVATION;
    Body b{};
    b.x = mk(d->x[0], d->x[1], d->x[2]);
    b.inv_mass = inv_mass;
    b.q = q_norm(Q4{{d->q[0], d->q[1], d->q[2], d->q[3]}});
    b.v = mk(d->v[0], d->v[1], d->v[2]);
    b.shape = d->shape;
    b.w = mk(d->w[0], d->w[1], d->w[2]);
    b.flags = flags;
    b.set_asleep(d->asleep && d->has_sleep);
    w->bodies.push_back(b);
    BodyConst c{};
    c.inv_inertia = inv_in;
    c.friction = d->friction;
    c.com_local = mk(d->com[0], d->com[1], d->com[2]);
    c.restitution = d->restitution;
    c.lin_damp = d->lin_damp;
    c.ang_damp = d->ang_damp;
    c.margin = d->margin;
    c.groups = d->groups;
    w->consts.push_back(c);
    w->prev.push_back(BodyPrev{b.x, 0.0f, b.q});
    w->objs.push_back(d->obj);
    w->scales.push_back(mk(d->scale[0], d->scale[1], d->scale[2]));
    w->sleep_lin.push_back(d->has_sleep ? d->sleep[0] : 0.0f);
    w->sleep_ang.push_back(d->has_sleep ? d->sleep[1] : 0.0f);
    w->timers.push_back(0.0f);
    *id = (uint32_t)w->bodies.size() - 1;
    return ORC_OK;
}

// World::step (world.rs:104-129): one frame of frame_dt seconds.
int orc_world_step(orc_world* w, float frame_dt) {
    ORC_CHECK(w, ORC_ERR_ARG, "orc_world_step: null");
    const float dt = frame_dt * w->time_scale;
    if (dt <= 0.0f || w->substeps == 0) return ORC_OK;
    collide(w, dt);
    const float h = dt / (float)w->substeps;
    const float g = len(w->gravity);
    std::vector<Body>& B = w->bodies;
    for (uint32_t sstep = 0; sstep < w->substeps; ++sstep) {
        presolve(B, w->contacts, w->vn0);
        for (size_t i = 0; i < B.size(); ++i) {  // integrate: half a kick, then the drift
            Body& b = B[i];
            w->prev[i] = BodyPrev{b.x, 0.0f, b.q};
            if (!b.moves()) continue;
            if (b.kind() == K_ACTIVE) b.v = add(b.v, scale(w->gravity, 0.5f * h));
            b.x = add(b.x, scale(b.v, h));
            b.q = q_integrate(b.q, b.w, h);
        }
        const float cap = rmax(w->max_recover * h / (float)std::max<uint32_t>(w->iterations, 1u), 0.0f);
        std::fill(w->push.begin(), w->push.end(), mk(0, 0, 0));
        std::fill(w->spin.begin(), w->spin.end(), mk(0, 0, 0));
        for (uint32_t it = 0; it < w->iterations; ++it) project(B, w->consts, w->prev, w->contacts, cap, w->push, w->spin);
        for (size_t i = 0; i < B.size(); ++i) {  // velocities from the poses (world.rs:183-193)
            Body& b = B[i];
            if (!b.dynamic()) continue;
            const BodyPrev& p = w->prev[i];
            V3 x_pred = add(p.x, scale(b.v, h));
            Q4 q_pred = q_integrate(p.q, b.w, h);
            b.v = add(b.v, scale(sub(sub(b.x, x_pred), w->push[i]), 1.0f / h));
            b.w = add(b.w, sub(q_delta_w(b.q, q_pred, h), scale(w->spin[i], 1.0f / h)));
        }
        velocity_pass(B, w->consts, w->contacts, w->vn0, h, g);
        for (size_t i = 0; i < B.size(); ++i) {  // the other half kick, then damping
            Body& b = B[i];
            if (!b.dynamic()) continue;
            const BodyConst& c = w->consts[i];
            b.v = add(b.v, scale(w->gravity, 0.5f * h));
            b.v = scale(b.v, std::pow(rmax(1.0f - c.lin_damp, 0.0f), h));
            b.w = scale(b.w, std::pow(rmax(1.0f - c.ang_damp, 0.0f), h));
        }
    }
    for (size_t i = 0; i < B.size(); ++i) {  // sleeping (world.rs:208-222)
        Body& b = B[i];
        if (!b.dynamic() || (b.flags & DEACTIVATION) == 0u) continue;
        const bool still = len(b.v) < w->sleep_lin[i] && len(b.w) < w->sleep_ang[i];
        w->timers[i] = still ? w->timers[i] + dt : 0.0f;
        if (w->timers[i] >= w->sleep_after) {
            b.set_asleep(true);
            b.v = mk(0, 0, 0);
            b.w = mk(0, 0, 0);
        }
    }
    return ORC_OK;
}

uint32_t orc_world_body_count(const orc_world* w) { return w ? (uint32_t)w->bodies.size() : 0u; }
uint32_t orc_world_contact_count(const orc_world* w) { return w ? (uint32_t)w->contacts.size() : 0u; }

// The bodies' 64-byte records (state.rs Body).
int orc_world_bodies(const orc_world* w, void* out) {
    ORC_CHECK(w && out, ORC_ERR_ARG, "orc_world_bodies: null");
    std::memcpy(out, w->bodies.data(), sizeof(Body) * w->bodies.size());
    return ORC_OK;
}

// World::poses (world.rs:225-228): each body's object matrix, m[col][row]
// (math.rs pose_matrix), 16 floats a body.
int orc_world_poses(const orc_world* w, float* out) {
    ORC_CHECK(w && out, ORC_ERR_ARG, "orc_world_poses: null");
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        const Body& b = w->bodies[i];
        const V3 s = w->scales[i];
        float* m = out + 16 * i;
        for (int k = 0; k < 3; ++k) {
            V3 e = mk(0, 0, 0);
            e[k] = s[k];
            V3 r = rotate(b.q, e);
            m[4 * k] = r[0], m[4 * k + 1] = r[1], m[4 * k + 2] = r[2], m[4 * k + 3] = 0.0f;
        }
        V3 t = sub(b.x, rotate(b.q, w->consts[i].com_local));
        m[12] = t[0], m[13] = t[1], m[14] = t[2], m[15] = 1.0f;
    }
    return ORC_OK;
}

// World::energy (world.rs:231-243): kinetic plus potential of the active bodies.
float orc_world_energy(const orc_world* w) {
    if (!w) return 0.0f;
    float total = 0.0f;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        const Body& b = w->bodies[i];
        const BodyConst& c = w->consts[i];
        if (!(b.inv_mass > 0.0f)) continue;
        const float m = 1.0f / b.inv_mass;
        V3 wb = unrotate(b.q, b.w);
        float rot = 0.0f;
        for (int k = 0; k < 3; ++k) rot = rot + (c.inv_inertia[k] > 0.0f ? 0.5f * wb[k] * wb[k] / c.inv_inertia[k] : 0.0f);
        total = total + ((0.5f * m * len2(b.v) + rot) - m * dot(w->gravity, b.x));
    }
    return total;
}

// The pairs the last broadphase found, sorted and deduplicated (broad.rs:73-91).
int orc_world_pair_list(const orc_world* w, uint32_t* out, uint32_t cap, uint32_t* n) {
    ORC_CHECK(w && n, ORC_ERR_ARG, "orc_world_pair_list: null");
    *n = (uint32_t)w->last_pairs.size();
    if (out) {
        ORC_CHECK(cap >= *n, ORC_ERR_SIZE, "orc_world_pair_list: buffer too small");
        for (uint32_t k = 0; k < *n; ++k) out[2 * k] = w->last_pairs[k].first, out[2 * k + 1] = w->last_pairs[k].second;
    }
    return ORC_OK;
}

}  // extern "C"

// CPU oracle: the LBVH builder, restated from phiray's `bvh::Bvh`.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates:
//   src/bvh/node.rs:78-105   Aabb::of_tri, union, center, morton/spread
//   src/bvh/mod.rs:28-43     code, delta
//   src/bvh/build.rs:28-194  Link::of, scene, codes, sort_codes, build,
//                            build_compact, assemble, refit
//   src/bvh/shape.rs:5-52    compact, reachable
//   src/bvh/sort.rs:119-150  the radix model (8 x 8-bit passes over (i, code))
// Parallel where the reference uses rayon (codes, sort, links); assemble and
// refit are serial, as they are there.
#include "oracle.h"

#include <algorithm>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

namespace {

// Rust's f32::min/max on non-NaN inputs (the only inputs the builder sees).
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return b > a ? b : a; }

template <class F>
void parallel_for(uint64_t n, int threads, F&& f) {
    if (threads <= 1 || n < 4096) {
        for (uint64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    uint64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        uint64_t a = per * t, b = std::min<uint64_t>(n, a + per);
        if (a >= b) break;
        pool.emplace_back([a, b, &f] { for (uint64_t i = a; i < b; ++i) f(i); });
    }
    for (auto& th : pool) th.join();
}

inline orc_aabb aabb_default() {
    orc_aabb b;
    // src/bvh/node.rs:67-71: (f32::MAX, f32::MIN)
    for (int k = 0; k < 3; ++k) { b.lo[k] = 3.40282347e38f; b.hi[k] = -3.40282347e38f; }
    b._a = b._b = 0.f;
    return b;
}

inline orc_aabb aabb_union(const orc_aabb& a, const orc_aabb& o) {
    // src/bvh/node.rs:82-84
    orc_aabb r;
    for (int k = 0; k < 3; ++k) { r.lo[k] = fmin_(a.lo[k], o.lo[k]); r.hi[k] = fmax_(a.hi[k], o.hi[k]); }
    r._a = r._b = 0.f;
    return r;
}

// src/bvh/node.rs:97-103
inline uint64_t spread(float v) {
    // Rust: (v.clamp(0.0, 1.0) * 1023.0) as u64 — a NaN stays NaN through
    // clamp and the saturating cast makes it 0.
    float c = v;
    if (c < 0.f) c = 0.f;
    if (c > 1.f) c = 1.f;
    uint64_t x = (c == c) ? (uint64_t)(c * 1023.0f) : 0;
    x = (x | (x << 16)) & 0x030000FFull;
    x = (x | (x << 8)) & 0x0300F00Full;
    x = (x | (x << 4)) & 0x030C30C3ull;
    return (x | (x << 2)) & 0x09249249ull;
}

inline uint32_t node_index(uint32_t i) { return i & ~ORC_LEAF; }
inline bool node_leaf(const orc_node& n) { return (n.lo.i & ORC_LEAF) != 0; }

}  // namespace

extern "C" {

int orc_hardware_threads(void) {
    unsigned t = std::thread::hardware_concurrency();
    return t ? (int)t : 1;
}

uint64_t orc_morton(const float p[3], uint32_t group) {
    // src/bvh/node.rs:96-105
    return ((uint64_t)group << 32) | (spread(p[0]) << 2) | (spread(p[1]) << 1) | spread(p[2]);
}

uint32_t orc_code(const float p[3]) { return (uint32_t)orc_morton(p, 0); }  // src/bvh/mod.rs:28-30

void orc_of_tri(const float a[3], const float b[3], const float c[3], orc_aabb* out) {
    // src/bvh/node.rs:78-80: a.min(b).min(c) per axis
    for (int k = 0; k < 3; ++k) {
        out->lo[k] = fmin_(fmin_(a[k], b[k]), c[k]);
        out->hi[k] = fmax_(fmax_(a[k], b[k]), c[k]);
    }
    out->_a = out->_b = 0.f;
}

void orc_scene(const orc_aabb* prims, uint32_t n, orc_aabb* out) {
    // src/bvh/build.rs:70-72: fold from Aabb::default()
    orc_aabb s = aabb_default();
    for (uint32_t i = 0; i < n; ++i) s = aabb_union(s, prims[i]);
    *out = s;
}

void orc_codes(const orc_aabb* prims, uint32_t n, uint32_t* codes, int threads) {
    // src/bvh/build.rs:75-85
    orc_aabb s;
    orc_scene(prims, n, &s);
    float inv[3];
    for (int k = 0; k < 3; ++k) {
        float e = s.hi[k] - s.lo[k];
        inv[k] = e > 0.f ? 1.0f / e : 0.f;
    }
    parallel_for(n, threads, [&](uint64_t i) {
        const orc_aabb& p = prims[i];
        float u[3];
        for (int k = 0; k < 3; ++k) {
            float c = 0.5f * (p.lo[k] + p.hi[k]);  // Aabb::center, src/bvh/node.rs:90-92
            u[k] = (c - s.lo[k]) * inv[k];
        }
        codes[i] = orc_code(u);
    });
}

void orc_sort_codes(const uint32_t* codes, uint32_t n, uint32_t* order, uint32_t* sorted, int threads) {
    // src/bvh/build.rs:87-92: indices sorted by (code, i) — a stable sort by code.
    std::vector<uint64_t> key(n), tmp(n);
    for (uint32_t i = 0; i < n; ++i) key[i] = ((uint64_t)codes[i] << 32) | i;
    // Chunked sort + pairwise merges: the same total order, run in parallel
    // like `par_sort_unstable_by_key`.
    int parts = (threads > 1 && n >= 65536) ? threads : 1;
    std::vector<uint64_t> bounds;
    for (int t = 0; t <= parts; ++t) bounds.push_back((uint64_t)n * t / parts);
    {
        std::vector<std::thread> pool;
        for (int t = 0; t < parts; ++t)
            pool.emplace_back([&, t] { std::sort(key.begin() + bounds[t], key.begin() + bounds[t + 1]); });
        for (auto& th : pool) th.join();
    }
    while (bounds.size() > 2) {
        std::vector<uint64_t> next{0};
        std::vector<std::thread> pool;
        size_t segs = bounds.size() - 1;
        for (size_t s = 0; s < segs; s += 2) {
            uint64_t a = bounds[s], m = bounds[s + 1];
            if (s + 1 < segs) {
                uint64_t b = bounds[s + 2];
                pool.emplace_back([&, a, m, b] {
                    std::merge(key.begin() + a, key.begin() + m, key.begin() + m, key.begin() + b, tmp.begin() + a);
                });
                next.push_back(b);
            } else {  // odd segment out: copied through
                pool.emplace_back([&, a, m] { std::copy(key.begin() + a, key.begin() + m, tmp.begin() + a); });
                next.push_back(m);
            }
        }
        for (auto& th : pool) th.join();
        key.swap(tmp);
        bounds.swap(next);
    }
    for (uint32_t k = 0; k < n; ++k) {
        order[k] = (uint32_t)key[k];
        if (sorted) sorted[k] = (uint32_t)(key[k] >> 32);
    }
}

void orc_radix_model(const uint32_t* codes, uint32_t n, uint32_t* order, uint32_t* sorted) {
    // src/bvh/sort.rs:119-150 (the model of Seed/Count/Offset/Scatter), with the
    // digit of src/bvh/radix.rs:69-72: 8 passes of 8 bits over the key
    // (lo = index, hi = code).
    std::vector<uint32_t> k0(n), k1(n), a0(n), a1(n);
    for (uint32_t i = 0; i < n; ++i) { k0[i] = i; k1[i] = codes[i]; }
    for (uint32_t d = 0; d < 8; ++d) {
        uint32_t shift = d * 8;
        auto digit = [&](uint32_t lo, uint32_t hi) { return shift < 32 ? (lo >> shift) & 0xFF : (hi >> (shift - 32)) & 0xFF; };
        uint32_t cursor[256] = {0};
        for (uint32_t i = 0; i < n; ++i) cursor[digit(k0[i], k1[i])]++;
        uint32_t running = 0;
        for (int b = 0; b < 256; ++b) { uint32_t c = cursor[b]; cursor[b] = running; running += c; }
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t b = digit(k0[i], k1[i]);
            a0[cursor[b]] = k0[i];
            a1[cursor[b]] = k1[i];
            cursor[b]++;
        }
        k0.swap(a0);
        k1.swap(a1);
    }
    for (uint32_t i = 0; i < n; ++i) { order[i] = k0[i]; if (sorted) sorted[i] = k1[i]; }
}

int32_t orc_delta(const uint32_t* codes, uint32_t n, int64_t i, int64_t j) {
    // src/bvh/mod.rs:33-43
    if (j < 0 || j >= (int64_t)n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a != b) return __builtin_clz(a ^ b);
    uint32_t x = (uint32_t)i ^ (uint32_t)j;
    return 32 + (x ? __builtin_clz(x) : 32);  // u32::leading_zeros(0) == 32
}

static orc_link link_of(const uint32_t* codes, uint32_t n32, int64_t i) {
    // src/bvh/build.rs:28-61 (Karras 2012, with the slot as tie-break)
    int64_t n = n32;
    auto delta = [&](int64_t a, int64_t b) { return orc_delta(codes, n32, a, b); };
    int64_t d = (delta(i, i + 1) - delta(i, i - 1) > 0) ? 1 : -1;
    int32_t dmin = delta(i, i - d);
    int64_t lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0, t = lmax / 2;
    while (t >= 1) {
        if (delta(i, i + (l + t) * d) > dmin) l += t;
        t /= 2;
    }
    int64_t j = i + l * d;
    int32_t dnode = delta(i, j);
    int64_t s = 0;
    t = l;
    for (;;) {
        t = (t + 1) >> 1;
        if (delta(i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int64_t g = i + s * d + std::min<int64_t>(d, 0);
    int64_t first = std::min(i, j), last = std::max(i, j);
    int64_t left = (first == g) ? n - 1 + g : g;
    int64_t right = (last == g + 1) ? n + g : g + 1;
    return orc_link{(uint32_t)left, (uint32_t)right, (uint32_t)first, (uint32_t)last};
}

void orc_links(const uint32_t* sorted, uint32_t n, orc_link* links, int threads) {
    if (n < 2) return;
    parallel_for(n - 1, threads, [&](uint64_t i) { links[i] = link_of(sorted, n, (int64_t)i); });
}

uint32_t orc_assemble(const orc_link* links, uint32_t n, uint32_t leaf, orc_node* nodes) {
    // src/bvh/build.rs:139-168
    if (n == 0) return 0;
    auto set = [](orc_node& nd, uint32_t a, uint32_t b) {
        std::memset(&nd, 0, sizeof nd);
        nd.lo.i = a;
        nd.hi.i = b;
    };
    for (uint32_t i = 0; i + 1 < n; ++i) set(nodes[i], links[i].left, links[i].right);
    for (uint32_t k = 0; k < n; ++k) set(nodes[n - 1 + k], k | ORC_LEAF, k + 1);
    uint32_t depth = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};
    while (!stack.empty()) {
        auto [i, d] = stack.back();
        stack.pop_back();
        depth = std::max(depth, d);
        if (i >= n - 1) continue;
        const orc_link& l = links[i];
        if (l.last - l.first < leaf || d + 1 >= ORC_STACK) {
            nodes[i].lo.i = l.first | ORC_LEAF;
            nodes[i].hi.i = l.last + 1;
            continue;
        }
        stack.push_back({l.left, d + 1});
        stack.push_back({l.right, d + 1});
    }
    return depth;
}

void orc_refit(orc_node* nodes, uint32_t count, const uint32_t* order, const orc_aabb* prims) {
    // src/bvh/build.rs:180-194: iterative post-order from the root.
    if (count == 0) return;
    std::vector<std::pair<uint32_t, bool>> stack{{0u, false}};
    while (!stack.empty()) {
        auto [i, done] = stack.back();
        stack.pop_back();
        orc_node nd = nodes[i];
        orc_aabb b;
        if (node_leaf(nd)) {
            b = aabb_default();
            for (uint32_t s = node_index(nd.lo.i); s < nd.hi.i; ++s) b = aabb_union(b, prims[order[s]]);
        } else if (done) {
            const orc_node& l = nodes[node_index(nd.lo.i)];
            const orc_node& r = nodes[node_index(nd.hi.i)];
            orc_aabb lb, rb;
            std::memcpy(lb.lo, l.lo.p, 12); std::memcpy(lb.hi, l.hi.p, 12);
            std::memcpy(rb.lo, r.lo.p, 12); std::memcpy(rb.hi, r.hi.p, 12);
            b = aabb_union(lb, rb);
        } else {
            stack.push_back({i, true});
            stack.push_back({node_index(nd.lo.i), false});
            stack.push_back({node_index(nd.hi.i), false});
            continue;
        }
        std::memcpy(nodes[i].lo.p, b.lo, 12);
        std::memcpy(nodes[i].hi.p, b.hi, 12);
    }
}

uint32_t orc_build(const orc_aabb* prims, uint32_t n, uint32_t leaf, orc_node* nodes, uint32_t* order, int threads) {
    // src/bvh/build.rs:94-100
    if (n == 0) return 0;
    std::vector<uint32_t> codes(n), sorted(n);
    orc_codes(prims, n, codes.data(), threads);
    orc_sort_codes(codes.data(), n, order, sorted.data(), threads);
    std::vector<orc_link> links(n > 1 ? n - 1 : 0);
    orc_links(sorted.data(), n, links.data(), threads);
    uint32_t depth = orc_assemble(links.data(), n, leaf, nodes);
    orc_refit(nodes, 2 * n - 1, order, prims);
    return depth;
}

uint32_t orc_build_compact(const orc_aabb* prims, uint32_t n, uint32_t leaf, orc_node* out, uint32_t* order, uint32_t* depth_out) {
    // src/bvh/build.rs:104-136: BFS numbering, children adjacent.
    *depth_out = 0;
    if (n == 0) return 0;
    std::vector<uint32_t> codes(n), sorted(n);
    orc_codes(prims, n, codes.data(), 1);
    orc_sort_codes(codes.data(), n, order, sorted.data(), 1);
    std::vector<orc_link> links(n > 1 ? n - 1 : 0);
    orc_links(sorted.data(), n, links.data(), 1);
    std::vector<orc_node> nodes(1);
    std::memset(nodes.data(), 0, sizeof(orc_node));
    struct Item { uint32_t i, slot, d; };
    std::deque<Item> queue{{0, 0, 0}};
    uint32_t depth = 0;
    auto set = [](orc_node& nd, uint32_t a, uint32_t b) { std::memset(&nd, 0, sizeof nd); nd.lo.i = a; nd.hi.i = b; };
    while (!queue.empty()) {
        Item it = queue.front();
        queue.pop_front();
        depth = std::max(depth, it.d);
        if (it.i >= n - 1) {
            uint32_t k = it.i - (n - 1);
            set(nodes[it.slot], k | ORC_LEAF, k + 1);
            continue;
        }
        const orc_link& l = links[it.i];
        if (l.last - l.first < leaf || it.d + 1 >= ORC_STACK) {
            set(nodes[it.slot], l.first | ORC_LEAF, l.last + 1);
            continue;
        }
        uint32_t at = (uint32_t)nodes.size();
        nodes.resize(at + 2);
        std::memset(&nodes[at], 0, 2 * sizeof(orc_node));
        set(nodes[it.slot], at, at + 1);
        queue.push_back({l.left, at, it.d + 1});
        queue.push_back({l.right, at + 1, it.d + 1});
    }
    orc_refit(nodes.data(), (uint32_t)nodes.size(), order, prims);
    std::memcpy(out, nodes.data(), nodes.size() * sizeof(orc_node));
    *depth_out = depth;
    return (uint32_t)nodes.size();
}

uint32_t orc_compact(orc_node* nodes, uint32_t count) {
    // src/bvh/shape.rs:5-37
    if (count == 0) return 0;
    std::vector<uint32_t> map(count, 0xffffffffu);
    std::vector<uint32_t> keep{0};
    map[0] = 0;
    for (size_t i = 0; i < keep.size(); ++i) {
        const orc_node& nd = nodes[keep[i]];
        if (!node_leaf(nd)) {
            for (uint32_t c : {node_index(nd.lo.i), node_index(nd.hi.i)}) {
                map[c] = (uint32_t)keep.size();
                keep.push_back(c);
            }
        }
    }
    std::vector<orc_node> out(keep.size());
    for (size_t k = 0; k < keep.size(); ++k) {
        orc_node nd = nodes[keep[k]];
        if (!node_leaf(nd)) {
            nd.lo.i = map[node_index(nd.lo.i)];
            nd.hi.i = map[node_index(nd.hi.i)];
        }
        out[k] = nd;
    }
    std::memcpy(nodes, out.data(), out.size() * sizeof(orc_node));
    return (uint32_t)out.size();
}

uint32_t orc_reachable(const orc_node* nodes, uint32_t count, uint32_t* out) {
    // src/bvh/shape.rs:39-51
    if (count == 0) return 0;
    uint32_t m = 0;
    std::vector<uint32_t> stack{0};
    while (!stack.empty()) {
        uint32_t i = stack.back();
        stack.pop_back();
        out[m++] = i;
        const orc_node& nd = nodes[i];
        if (!node_leaf(nd)) {
            stack.push_back(node_index(nd.lo.i));
            stack.push_back(node_index(nd.hi.i));
        }
    }
    return m;
}

}  // extern "C"

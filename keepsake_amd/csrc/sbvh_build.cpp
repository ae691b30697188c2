// Host SBVH builder. The split search restates the reference builder (cuda/sbvh.cu:81-440):
// binned SAH object split along the largest centroid axis, spatial split candidates when the two
// object-split children overlap by more than alpha * root area, reference unsplitting, leaf when the
// SAH says so and the node holds at most maxPrimsInNode references. The result is flattened into a
// BVH2 whose nodes carry both children's boxes (64 bytes) and whose leaves point into a triangle
// array stored in leaf order (48 bytes per reference), the layout the gfx950 traversal kernels read.
#include "ks_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <limits>
#include <stdexcept>
#include <thread>

namespace ksa {
namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
    float lo[3] = {kInf, kInf, kInf};
    float hi[3] = {-kInf, -kInf, -kInf};
    void expand(const float p[3]) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], p[i]);
            hi[i] = std::max(hi[i], p[i]);
        }
    }
    void expand(const Box &b) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], b.lo[i]);
            hi[i] = std::max(hi[i], b.hi[i]);
        }
    }
    bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
    float area() const {
        float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
        return 2.0f * (e0 * e1 + e1 * e2 + e0 * e2);
    }
    float center(int d) const { return 0.5f * (lo[d] + hi[d]); }
    float extent(int d) const { return hi[d] - lo[d]; }
    int largest_axis() const {
        float e0 = extent(0), e1 = extent(1), e2 = extent(2);
        if (e0 >= e1 && e0 >= e2) return 0;
        if (e1 >= e0 && e1 >= e2) return 1;
        return 2;
    }
    // AABB3::offset along one axis (cuda/aabb.cuh:40-47)
    float offset(float p, int d) const {
        float e = extent(d);
        if (e == 0.0f) return 0.0f;
        return (p - lo[d]) / e;
    }
};

inline Box join(const Box &a, const Box &b) {
    Box r = a;
    r.expand(b);
    return r;
}
inline Box intersect(const Box &a, const Box &b) {
    Box r;
    for (int i = 0; i < 3; ++i) {
        r.lo[i] = std::max(a.lo[i], b.lo[i]);
        r.hi[i] = std::min(a.hi[i], b.hi[i]);
    }
    return r;
}
inline float lerp_exact(float a, float b, float t) { return t == 1.0f ? b : a + t * (b - a); }

struct Ref {
    uint32_t tri;
    Box box; // partial bound after spatial splits
};

struct BuildNode {
    Box bound;             // node bound
    Box cbound[2];         // children bounds as the split reported them
    uint32_t child[2] = {0, 0};
    uint32_t axis = 0;
    uint32_t first = 0, count = 0; // leaf references in `ordered`
    bool leaf = false;
};

struct Split {
    Box left, right;
    float cost = kInf;
    uint32_t bucket = 0;
};

struct Builder {
    ks_bvh_options opt;
    float root_area = 0.0f;
    std::vector<BuildNode> nodes;
    std::vector<uint32_t> ordered;
    std::mutex mu;

    // estimateObjectSplit (cuda/sbvh.cu:81-139)
    Split object_split(const std::vector<Ref> &refs, const Box &bounds, const Box &cb, int dim, bool &skip_spatial) {
        uint32_t B = opt.bucket_count;
        std::vector<uint32_t> cnt(B, 0);
        std::vector<Box> bb(B);
        for (const Ref &r : refs) {
            uint32_t b = (uint32_t)std::floor(B * cb.offset(r.box.center(dim), dim));
            b = std::min(b, B - 1);
            ++cnt[b];
            bb[b].expand(r.box);
        }
        // prefix/suffix sweeps give the same bounds and counts as the reference double loop
        std::vector<Box> lb(B), rb(B);
        std::vector<uint32_t> lc(B), rc(B);
        Box acc;
        uint32_t c = 0;
        for (uint32_t i = 0; i < B; ++i) {
            acc.expand(bb[i]);
            c += cnt[i];
            lb[i] = acc;
            lc[i] = c;
        }
        acc = Box();
        c = 0;
        for (int i = (int)B - 1; i >= 0; --i) {
            acc.expand(bb[i]);
            c += cnt[i];
            rb[i] = acc;
            rc[i] = c;
        }
        Split best;
        float inv_area = 1.0f / bounds.area();
        for (uint32_t i = 0; i + 1 < B; ++i) {
            float cost = opt.trav_weight + (lc[i] * lb[i].area() + rc[i + 1] * rb[i + 1].area()) * inv_area;
            if (i == 0 || cost < best.cost) {
                best.cost = cost;
                best.bucket = i;
                best.left = lb[i];
                best.right = rb[i + 1];
            }
        }
        Box is = intersect(best.left, best.right);
        float isa = is.empty() ? 0.0f : is.area();
        skip_spatial = isa / root_area <= opt.alpha;
        return best;
    }

    // estimateSpatialSplit (cuda/sbvh.cu:141-207)
    Split spatial_split(const std::vector<Ref> &refs, const Box &bounds, int dim) {
        uint32_t B = opt.bucket_count;
        std::vector<uint32_t> enter(B, 0), exitc(B, 0);
        std::vector<Box> bb(B);
        std::vector<float> planes(B + 1);
        for (uint32_t b = 0; b <= B; ++b) planes[b] = lerp_exact(bounds.lo[dim], bounds.hi[dim], (float)b / (float)B);
        for (const Ref &r : refs) {
            uint32_t en = B - 1, ex = 0;
            for (uint32_t b = 0; b < B; ++b) {
                if (r.box.lo[dim] > planes[b + 1]) continue;
                if (r.box.hi[dim] < planes[b]) break;
                ex = std::max(ex, b);
                en = std::min(en, b);
                Box ch = r.box;
                ch.lo[dim] = std::max(ch.lo[dim], planes[b]);
                ch.hi[dim] = std::min(ch.hi[dim], planes[b + 1]);
                bb[b].expand(ch);
            }
            if (en > ex) en = ex; // numerically empty overlap: keep the reference in one bucket
            ++enter[en];
            ++exitc[ex];
        }
        std::vector<Box> lb(B), rb(B);
        std::vector<uint32_t> lc(B), rc(B);
        Box acc;
        uint32_t c = 0;
        for (uint32_t i = 0; i < B; ++i) {
            acc.expand(bb[i]);
            c += enter[i];
            lb[i] = acc;
            lc[i] = c;
        }
        acc = Box();
        c = 0;
        for (int i = (int)B - 1; i >= 0; --i) {
            acc.expand(bb[i]);
            c += exitc[i];
            rb[i] = acc;
            rc[i] = c;
        }
        Split best;
        float inv_area = 1.0f / bounds.area();
        for (uint32_t i = 0; i + 1 < B; ++i) {
            float cost = opt.trav_weight + (lc[i] * lb[i].area() + rc[i + 1] * rb[i + 1].area()) * inv_area;
            if (i == 0 || cost < best.cost) {
                best.cost = cost;
                best.bucket = i;
                best.left = lb[i];
                best.right = rb[i + 1];
            }
        }
        return best;
    }

    // doObjectSplit (cuda/sbvh.cu:215-231)
    uint32_t do_object_split(std::vector<Ref> &refs, const Split &est, const Box &cb, int dim) {
        uint32_t B = opt.bucket_count;
        auto mid = std::partition(refs.begin(), refs.end(), [&](const Ref &r) {
            uint32_t b = (uint32_t)std::floor(B * cb.offset(r.box.center(dim), dim));
            b = std::min(b, B - 1);
            return b <= est.bucket;
        });
        return (uint32_t)(mid - refs.begin());
    }

    // doSpatialSplit with reference unsplitting (cuda/sbvh.cu:233-298)
    uint32_t do_spatial_split(std::vector<Ref> &refs, const Split &est, const Box &bounds, int dim, Box &al,
                              Box &ar) {
        float clip = lerp_exact(bounds.lo[dim], bounds.hi[dim], (float)(est.bucket + 1) / (float)opt.bucket_count);
        int left_end = 0, right_start = (int)refs.size();
        for (int i = left_end; i < right_start; i++) {
            if (refs[i].box.hi[dim] <= clip) {
                al.expand(refs[i].box);
                std::swap(refs[i], refs[left_end++]);
            } else if (refs[i].box.lo[dim] >= clip) {
                ar.expand(refs[i].box);
                std::swap(refs[i--], refs[--right_start]);
            }
        }
        std::vector<Ref> extra;
        int n1 = right_start;
        int n2 = (int)refs.size() - left_end;
        float full = est.left.area() * n1 + est.right.area() * n2;
        while (left_end < right_start) {
            Ref &cur = refs[left_end];
            float move_left = join(est.left, cur.box).area() * n1 + est.right.area() * (n2 - 1);
            float move_right = est.left.area() * (n1 - 1) + join(est.right, cur.box).area() * n2;
            float m = std::min(full, std::min(move_left, move_right));
            if (m == move_left) {
                al.expand(cur.box);
                ++left_end;
            } else if (m == move_right) {
                ar.expand(cur.box);
                std::swap(cur, refs[--right_start]);
            } else {
                Ref nr = cur;
                nr.box.lo[dim] = clip;
                cur.box.hi[dim] = clip;
                al.expand(cur.box);
                ar.expand(nr.box);
                ++left_end;
                extra.push_back(nr);
            }
        }
        refs.insert(refs.end(), extra.begin(), extra.end());
        return (uint32_t)right_start;
    }

    uint32_t alloc_node() {
        std::lock_guard<std::mutex> g(mu);
        nodes.emplace_back();
        return (uint32_t)nodes.size() - 1;
    }

    // recursiveBuild (cuda/sbvh.cu:300-369); `ordered` ranges are assigned in a second pass so the
    // top levels may build in parallel.
    uint32_t build(std::vector<Ref> &refs, int depth_budget, std::vector<uint32_t> &leaf_refs_out) {
        uint32_t ni = alloc_node();
        Box bounds;
        for (const Ref &r : refs) bounds.expand(r.box);
        auto as_leaf = [&]() {
            BuildNode n;
            n.bound = bounds;
            n.leaf = true;
            n.count = (uint32_t)refs.size();
            n.first = (uint32_t)leaf_refs_out.size();
            for (const Ref &r : refs) leaf_refs_out.push_back(r.tri);
            std::lock_guard<std::mutex> g(mu);
            nodes[ni] = n;
            return ni;
        };
        uint32_t n = (uint32_t)refs.size();
        if (n == 1) return as_leaf();
        Box cb;
        for (const Ref &r : refs) {
            float c[3] = {r.box.center(0), r.box.center(1), r.box.center(2)};
            cb.expand(c);
        }
        int dim = cb.largest_axis();
        if (cb.hi[dim] == cb.lo[dim]) return as_leaf();
        float leaf_cost = (float)n;
        bool skip_spatial = true;
        Split os = object_split(refs, bounds, cb, dim, skip_spatial);
        Split ss;
        if (!skip_spatial) ss = spatial_split(refs, bounds, dim);
        if (n <= opt.max_prims_in_node && leaf_cost <= os.cost && (skip_spatial || leaf_cost <= ss.cost))
            return as_leaf();
        uint32_t mid;
        Box lbound, rbound;
        if (skip_spatial || os.cost <= ss.cost) {
            mid = do_object_split(refs, os, cb, dim);
            lbound = os.left;
            rbound = os.right;
        } else {
            mid = do_spatial_split(refs, ss, bounds, dim, lbound, rbound);
        }
        if (mid == 0 || mid >= refs.size()) {
            // cannot happen for the reference's split rules; guard against float corner cases
            std::nth_element(refs.begin(), refs.begin() + n / 2, refs.end(),
                             [&](const Ref &a, const Ref &b) { return a.box.center(dim) < b.box.center(dim); });
            mid = n / 2;
            lbound = Box();
            rbound = Box();
            for (uint32_t i = 0; i < mid; ++i) lbound.expand(refs[i].box);
            for (uint32_t i = mid; i < refs.size(); ++i) rbound.expand(refs[i].box);
        }
        std::vector<Ref> right(refs.begin() + mid, refs.end());
        refs.resize(mid);
        refs.shrink_to_fit();
        uint32_t l, r;
        if (depth_budget > 0 && n > 4096) {
            std::vector<uint32_t> lrefs, rrefs;
            auto fut = std::async(std::launch::async, [&]() { return build(right, depth_budget - 1, rrefs); });
            l = build(refs, depth_budget - 1, lrefs);
            r = fut.get();
            // stitch the leaf reference lists: right subtree leaf offsets shift by the left list size
            uint32_t base = (uint32_t)leaf_refs_out.size();
            uint32_t lsize = (uint32_t)lrefs.size();
            leaf_refs_out.insert(leaf_refs_out.end(), lrefs.begin(), lrefs.end());
            leaf_refs_out.insert(leaf_refs_out.end(), rrefs.begin(), rrefs.end());
            shift_leaves(l, base);
            shift_leaves(r, base + lsize);
        } else {
            l = build(refs, 0, leaf_refs_out);
            r = build(right, 0, leaf_refs_out);
        }
        BuildNode nd;
        nd.bound = join(lbound, rbound);
        nd.cbound[0] = lbound;
        nd.cbound[1] = rbound;
        nd.child[0] = l;
        nd.child[1] = r;
        nd.axis = (uint32_t)dim;
        std::lock_guard<std::mutex> g(mu);
        nodes[ni] = nd;
        return ni;
    }

    void shift_leaves(uint32_t root, uint32_t base) {
        std::vector<uint32_t> st{root};
        while (!st.empty()) {
            uint32_t i = st.back();
            st.pop_back();
            BuildNode *n;
            {
                std::lock_guard<std::mutex> g(mu);
                n = &nodes[i];
                if (n->leaf) {
                    n->first += base;
                    continue;
                }
                st.push_back(n->child[0]);
                st.push_back(n->child[1]);
            }
        }
    }
};

struct Flattener {
    const Builder &b;
    const float *verts;
    const uint32_t *idx;
    std::vector<GpuNode> &out;
    std::vector<GpuTri> &tris;
    uint32_t max_depth = 0;

    GpuTri make_tri(uint32_t t) const {
        GpuTri g;
        const float *a = verts + 3 * idx[3 * t], *bb = verts + 3 * idx[3 * t + 1], *c = verts + 3 * idx[3 * t + 2];
        for (int k = 0; k < 3; ++k) {
            g.v0[k] = a[k];
            g.v1[k] = bb[k];
            g.v2[k] = c[k];
        }
        g.tri = t;
        g.v1[3] = 0.0f;
        g.v2[3] = 0.0f;
        return g;
    }

    Box tri_box(uint32_t t) const {
        Box bx;
        for (int k = 0; k < 3; ++k) bx.expand(verts + 3 * idx[3 * t + k]);
        return bx;
    }

    static void put_box(GpuNode &n, int slot, const Box &bx) {
        float *xy = n.f + 4 * slot;
        xy[0] = bx.lo[0];
        xy[1] = bx.hi[0];
        xy[2] = bx.lo[1];
        xy[3] = bx.hi[1];
        n.f[8 + 2 * slot] = bx.lo[2];
        n.f[9 + 2 * slot] = bx.hi[2];
    }

    // Emits references [first, first+count) of a leaf; returns the child pointer. Leaves larger than
    // kMaxLeafRefs become a balanced subtree (only degenerate inputs produce them).
    uint32_t emit_leaf(const std::vector<uint32_t> &refs, uint32_t first, uint32_t count, int depth) {
        if (count <= kMaxLeafRefs) {
            uint32_t start = (uint32_t)tris.size();
            if (start + count > kMaxRefs) throw std::runtime_error("too many triangle references for the BVH layout");
            for (uint32_t i = 0; i < count; ++i) tris.push_back(make_tri(refs[first + i]));
            max_depth = std::max<uint32_t>(max_depth, depth);
            return kLeafBit | (start << 4) | (count - 1);
        }
        uint32_t ni = (uint32_t)out.size();
        out.emplace_back();
        uint32_t h = count / 2;
        Box lb, rb;
        for (uint32_t i = 0; i < h; ++i) lb.expand(tri_box(refs[first + i]));
        for (uint32_t i = h; i < count; ++i) rb.expand(tri_box(refs[first + i]));
        uint32_t c0 = emit_leaf(refs, first, h, depth + 1);
        uint32_t c1 = emit_leaf(refs, first + h, count - h, depth + 1);
        GpuNode &n = out[ni];
        put_box(n, 0, lb);
        put_box(n, 1, rb);
        n.child[0] = c0;
        n.child[1] = c1;
        n.pad[0] = n.pad[1] = 0;
        return ni;
    }

    uint32_t emit(uint32_t bi, const std::vector<uint32_t> &refs, int depth) {
        const BuildNode &bn = b.nodes[bi];
        if (bn.leaf) return emit_leaf(refs, bn.first, bn.count, depth);
        uint32_t ni = (uint32_t)out.size();
        out.emplace_back();
        uint32_t c0 = emit(bn.child[0], refs, depth + 1);
        uint32_t c1 = emit(bn.child[1], refs, depth + 1);
        GpuNode &n = out[ni];
        // the stored child bounds are the tight bounds of each subtree (the split's reported bounds
        // can be looser after reference unsplitting)
        put_box(n, 0, b.nodes[bn.child[0]].bound);
        put_box(n, 1, b.nodes[bn.child[1]].bound);
        n.child[0] = c0;
        n.child[1] = c1;
        n.pad[0] = bn.axis; // SBVHNode::splitAxis (cuda/sbvh.cuh:22), for the reference-order traversal count
        n.pad[1] = 0;
        return ni;
    }
};

} // namespace

BvhResult build_sbvh(const float *vertices, const uint32_t *indices, uint32_t n_tris, const ks_bvh_options &o) {
    ks_bvh_options opt = o;
    if (opt.max_prims_in_node == 0) opt.max_prims_in_node = 4;
    if (opt.bucket_count == 0) opt.bucket_count = 12;
    if (opt.trav_weight == 0.0f) opt.trav_weight = 1.0f;
    if (opt.alpha == 0.0f) opt.alpha = 1e-5f;
    // clamps of SBVH::build (cuda/sbvh.cu:411-414)
    opt.max_prims_in_node = std::min(opt.max_prims_in_node, 8u);
    opt.bucket_count = std::max(2u, std::min(opt.bucket_count, 16u));
    opt.trav_weight = std::min(std::max(opt.trav_weight, 0.0f), 10.0f);
    opt.alpha = std::min(std::max(opt.alpha, 0.0f), 1.0f);

    BvhResult res;
    Builder b;
    b.opt = opt;
    std::vector<Ref> refs;
    refs.reserve(n_tris);
    Box scene;
    for (uint32_t t = 0; t < n_tris; ++t) {
        Ref r;
        r.tri = t;
        for (int k = 0; k < 3; ++k) r.box.expand(vertices + 3 * indices[3 * t + k]);
        // degenerate (zero-area in every axis or non-finite) triangles still get a reference; the
        // watertight test rejects them
        refs.push_back(r);
        scene.expand(r.box);
    }
    for (int k = 0; k < 3; ++k) {
        res.lo[k] = n_tris ? scene.lo[k] : 0.0f;
        res.hi[k] = n_tris ? scene.hi[k] : 0.0f;
    }
    if (n_tris == 0) {
        GpuNode n{};
        Box e;
        Flattener::put_box(n, 0, e);
        Flattener::put_box(n, 1, e);
        n.child[0] = n.child[1] = kLeafBit; // never visited: empty boxes
        res.nodes.push_back(n);
        return res;
    }
    b.root_area = scene.area();
    if (b.root_area <= 0.0f) b.root_area = 1.0f;
    b.nodes.reserve(2 * (size_t)n_tris);
    int threads = opt.threads > 0 ? opt.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    int budget = 0;
    while ((1 << budget) < threads && budget < 8) ++budget;
    std::vector<uint32_t> leaf_refs;
    leaf_refs.reserve(n_tris + n_tris / 4);
    uint32_t root = b.build(refs, budget, leaf_refs);

    Flattener f{b, vertices, indices, res.nodes, res.tris};
    res.nodes.reserve(b.nodes.size());
    res.tris.reserve(leaf_refs.size());
    const BuildNode &rn = b.nodes[root];
    if (rn.leaf) {
        // a single-leaf tree still gets an interior root so traversal always starts at node 0
        res.nodes.emplace_back();
        uint32_t c = f.emit_leaf(leaf_refs, rn.first, rn.count, 1);
        GpuNode &n = res.nodes[0];
        Flattener::put_box(n, 0, rn.bound);
        Flattener::put_box(n, 1, Box());
        n.child[0] = c;
        n.child[1] = kLeafBit; // empty box: never entered
        n.pad[0] = n.pad[1] = 0;
        if (res.nodes[0].child[0] != c) res.nodes[0].child[0] = c;
    } else {
        f.emit(root, leaf_refs, 0);
    }
    res.max_depth = f.max_depth;
    res.nodes4 = collapse_bvh4(res.nodes, res.max_depth4);
    return res;
}

namespace {

struct Child2 {
    float lo[3], hi[3];
    uint32_t ptr; // BVH2 child pointer
    bool interior() const { return (ptr & kLeafBit) == 0; }
    float area() const {
        float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        if (!(e[0] >= 0.0f && e[1] >= 0.0f && e[2] >= 0.0f)) return -1.0f; // empty box
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
    }
};

Child2 child_of(const GpuNode &n, int slot) {
    Child2 c;
    c.lo[0] = n.f[4 * slot];
    c.hi[0] = n.f[4 * slot + 1];
    c.lo[1] = n.f[4 * slot + 2];
    c.hi[1] = n.f[4 * slot + 3];
    c.lo[2] = n.f[8 + 2 * slot];
    c.hi[2] = n.f[9 + 2 * slot];
    c.ptr = n.child[slot];
    return c;
}

struct Collapser {
    const std::vector<GpuNode> &in;
    std::vector<GpuNode4> &out;
    uint32_t max_depth = 0;

    uint32_t emit(uint32_t n2, uint32_t depth) {
        std::vector<Child2> ch{child_of(in[n2], 0), child_of(in[n2], 1)};
        while (ch.size() < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (size_t i = 0; i < ch.size(); ++i)
                if (ch[i].interior() && ch[i].area() > best_area) {
                    best = (int)i;
                    best_area = ch[i].area();
                }
            if (best < 0) break;
            const GpuNode &g = in[ch[best].ptr];
            ch[best] = child_of(g, 0);
            ch.insert(ch.begin() + best + 1, child_of(g, 1));
        }
        uint32_t ni = (uint32_t)out.size();
        out.emplace_back();
        uint32_t ptr[4];
        for (size_t i = 0; i < ch.size(); ++i) {
            if (ch[i].interior()) {
                ptr[i] = emit(ch[i].ptr, depth + 1);
            } else {
                ptr[i] = ch[i].ptr;
                max_depth = std::max(max_depth, depth + 1);
            }
        }
        GpuNode4 &n = out[ni];
        const float inf = std::numeric_limits<float>::infinity();
        for (int i = 0; i < 4; ++i) {
            bool used = i < (int)ch.size();
            n.lox[i] = used ? ch[i].lo[0] : inf;
            n.hix[i] = used ? ch[i].hi[0] : -inf;
            n.loy[i] = used ? ch[i].lo[1] : inf;
            n.hiy[i] = used ? ch[i].hi[1] : -inf;
            n.loz[i] = used ? ch[i].lo[2] : inf;
            n.hiz[i] = used ? ch[i].hi[2] : -inf;
            n.child[i] = used ? ptr[i] : kLeafBit;
            n.pad[i] = 0;
        }
        return ni;
    }
};

} // namespace

std::vector<GpuNode4> collapse_bvh4(const std::vector<GpuNode> &nodes, uint32_t &max_depth4) {
    std::vector<GpuNode4> out;
    out.reserve(nodes.size() / 2 + 1);
    Collapser c{nodes, out};
    if (!nodes.empty()) c.emit(0, 0);
    max_depth4 = c.max_depth;
    return out;
}

} // namespace ksa

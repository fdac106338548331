// Host BVH builder: binned-SAH binary tree, collapsed to the 4-wide node the kernels traverse.
//
// Replaces the acceleration structure Embree builds in rtcCommit (rt/Embree.cpp:55-58 via
// Scene::end, scene/Scene.cpp:122-142). Child boxes are the fp32 min/max of their vertices,
// widened by a padding of 2^-18 of the largest coordinate magnitude: the kernels' fma slab test
// shifts a plane by at most 3 |o| * 2^-24, and the exit distances are scaled by 1 + 2*gamma(3)
// (Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013), so no box the exact ray crosses is missed
// for ray origins within 20 times the largest coordinate magnitude.
#include "rl_bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace rl {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; k++) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Node2 {
    Box box;
    int32_t left = -1, right = -1;  // internal when left >= 0
    uint32_t first = 0, count = 0;
};

constexpr int kBins = 32;
constexpr uint32_t kLeafTarget = 4;  // binary leaves hold at most this many triangles

class Builder2 {
public:
    Builder2(const float* tri9, uint32_t n) : n_(n) {
        tbox_.resize(n);
        cen_.resize(n * 3);
        order_.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            Box b;
            b.reset();
            b.grow(tri9 + 9 * i);
            b.grow(tri9 + 9 * i + 3);
            b.grow(tri9 + 9 * i + 6);
            tbox_[i] = b;
            for (int k = 0; k < 3; k++) cen_[3 * i + k] = 0.5f * (b.lo[k] + b.hi[k]);
            order_[i] = i;
        }
    }

    void build() {
        nodes_.reserve(2 * (size_t)n_ / kLeafTarget + 16);
        nodes_.emplace_back();
        struct Item { int32_t node; uint32_t first, count; };
        std::vector<Item> work;
        work.push_back({0, 0, n_});
        while (!work.empty()) {
            Item it = work.back();
            work.pop_back();
            Box box, cbox;
            box.reset();
            cbox.reset();
            for (uint32_t i = it.first; i < it.first + it.count; i++) {
                box.grow(tbox_[order_[i]]);
                cbox.grow(&cen_[3 * order_[i]]);
            }
            nodes_[it.node].box = box;
            uint32_t mid = 0;
            if (it.count <= 1 || !split(it.first, it.count, box, cbox, mid)) {
                nodes_[it.node].first = it.first;
                nodes_[it.node].count = it.count;
                continue;
            }
            int32_t l = (int32_t)nodes_.size();
            nodes_.emplace_back();
            nodes_.emplace_back();
            nodes_[it.node].left = l;
            nodes_[it.node].right = l + 1;
            work.push_back({l + 1, mid, it.first + it.count - mid});
            work.push_back({l, it.first, mid - it.first});
        }
    }

    std::vector<Node2> nodes_;
    std::vector<uint32_t> order_;

private:
    // Chooses a split of [first, first+count) by binned SAH; returns false to make a leaf.
    bool split(uint32_t first, uint32_t count, const Box& box, const Box& cbox, uint32_t& mid) {
        int axis = 0;
        float ext[3];
        for (int k = 0; k < 3; k++) ext[k] = cbox.hi[k] - cbox.lo[k];
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        uint32_t* ord = order_.data() + first;
        if (!(ext[axis] > 0.0f)) {
            // All centroids coincide: split by index once the leaf would be too large.
            if (count <= RL_LEAF_MAX) return false;
            mid = first + count / 2;
            return true;
        }
        const float lo = cbox.lo[axis];
        const float scale = (float)kBins / ext[axis];
        auto bin_of = [&](uint32_t t) {
            int b = (int)((cen_[3 * t + axis] - lo) * scale);
            return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
        };
        Box bb[kBins];
        uint32_t bc[kBins] = {0};
        for (int b = 0; b < kBins; b++) bb[b].reset();
        for (uint32_t i = 0; i < count; i++) {
            int b = bin_of(ord[i]);
            bb[b].grow(tbox_[ord[i]]);
            bc[b]++;
        }
        double right_area[kBins];
        uint32_t right_cnt[kBins];
        Box acc;
        acc.reset();
        uint32_t cnt = 0;
        for (int b = kBins - 1; b > 0; b--) {
            acc.grow(bb[b]);
            cnt += bc[b];
            right_area[b] = acc.area();
            right_cnt[b] = cnt;
        }
        acc.reset();
        cnt = 0;
        double best = DBL_MAX;
        int best_b = -1;
        for (int b = 1; b < kBins; b++) {
            acc.grow(bb[b - 1]);
            cnt += bc[b - 1];
            if (cnt == 0 || right_cnt[b] == 0) continue;
            double c = acc.area() * cnt + right_area[b] * right_cnt[b];
            if (c < best) { best = c; best_b = b; }
        }
        double leaf_cost = box.area() * count;
        // A traversal step costs half a triangle test, compared against the leaf alternative.
        // Measured on the city (AO, direct light) and the heightfield final gather with at most
        // 2/3/4/6/8 triangles per binary leaf and step costs 0.5-4: the kernels wait on fetches, so
        // fatter leaves lose (4 triangles, cost 2: AO +7 %, FG +2 %; 8 and 4: AO +78 %) and thin
        // ones are level; 4 triangles at cost 0.5 measured AO -0.7 %, direct -2.4 %, FG level.
        double split_cost = 0.5 * box.area() + best;
        if (best_b < 0) {
            if (count <= RL_LEAF_MAX) return false;
            mid = first + count / 2;
            std::nth_element(ord, ord + count / 2, ord + count,
                             [&](uint32_t a, uint32_t b) { return cen_[3 * a + axis] < cen_[3 * b + axis]; });
            return true;
        }
        if (count <= kLeafTarget && split_cost >= leaf_cost) return false;
        uint32_t* p = std::partition(ord, ord + count, [&](uint32_t t) { return bin_of(t) < best_b; });
        mid = first + (uint32_t)(p - ord);
        if (mid == first || mid == first + count) {
            mid = first + count / 2;
        }
        return true;
    }

    uint32_t n_;
    std::vector<Box> tbox_;
    std::vector<float> cen_;
};

class Collapser {
public:
    Collapser(const std::vector<Node2>& n2, float pad) : n2_(n2), pad_(pad) {}

    // Builds the 4-wide tree rooted at binary node r; the root is always an internal node.
    void run() {
        out_.reserve(n2_.size() / 2 + 4);
        out_.emplace_back();
        fill(0, 0);
    }

    std::vector<Bvh4Node> out_;
    int stack_bound_ = 0;
    int depth_ = 0;

private:
    static double area(const Node2& n) { return n.box.area(); }

    // Fills out_[dst] from binary node src; returns the traversal-stack bound below dst.
    int fill(uint32_t dst, int32_t src) {
        std::vector<int32_t> kids;
        const Node2& s = n2_[src];
        if (s.left < 0) {
            kids.push_back(src);  // leaf root: wrap it in a one-child node
        } else {
            kids.push_back(s.left);
            kids.push_back(s.right);
        }
        while (kids.size() < 4) {
            int best = -1;
            double ba = -1.0;
            for (size_t i = 0; i < kids.size(); i++) {
                const Node2& k = n2_[kids[i]];
                if (k.left >= 0 && area(k) > ba) { ba = area(k); best = (int)i; }
            }
            if (best < 0) break;
            int32_t open = kids[best];
            kids[best] = n2_[open].left;
            kids.push_back(n2_[open].right);
        }
        Bvh4Node node;
        std::memset(&node, 0, sizeof(node));
        int child_bound = 0;
        std::vector<std::pair<int, int32_t>> recurse;
        for (int i = 0; i < 4; i++) {
            if (i >= (int)kids.size()) {
                // Empty slot: an inverted box (lo = +1e30, hi = -1e30). For either sign of a
                // direction component its entry distance is positive and huge (or +inf) and its
                // exit distance negative, so the slab test never accepts it.
                node.lox[i] = node.loy[i] = node.loz[i] = 1e30f;
                node.hix[i] = node.hiy[i] = node.hiz[i] = -1e30f;
                node.child[i] = RL_EMPTY_CHILD;
                continue;
            }
            const Node2& k = n2_[kids[i]];
            node.lox[i] = k.box.lo[0] - pad_; node.hix[i] = k.box.hi[0] + pad_;
            node.loy[i] = k.box.lo[1] - pad_; node.hiy[i] = k.box.hi[1] + pad_;
            node.loz[i] = k.box.lo[2] - pad_; node.hiz[i] = k.box.hi[2] + pad_;
            if (k.left < 0) {
                node.child[i] = RL_LEAF_BIT | ((k.count - 1) << 27) | k.first;
            } else {
                recurse.push_back({i, kids[i]});
            }
        }
        out_[dst] = node;
        for (auto& r : recurse) {
            uint32_t idx = (uint32_t)out_.size();
            out_.emplace_back();
            out_[dst].child[r.first] = idx;
            depth_cur_++;
            depth_ = std::max(depth_, depth_cur_);
            int b = fill(idx, r.second);
            depth_cur_--;
            child_bound = std::max(child_bound, b);
        }
        int here = (int)kids.size() - 1 + child_bound;
        if (dst == 0) stack_bound_ = here;
        return here;
    }

    const std::vector<Node2>& n2_;
    float pad_;
    int depth_cur_ = 1;
};

}  // namespace

int build_bvh4(const float* tri9, uint32_t num_tris, Bvh4Build& out) {
    if (num_tris == 0) return -1;
    if (num_tris > RL_LEAF_FIRST_MASK) return -2;
    Builder2 b2(tri9, num_tris);
    b2.build();
    for (const Node2& n : b2.nodes_) {
        if (n.left < 0 && n.count > RL_LEAF_MAX) return -3;
    }
    // Box padding for the fma slab test: 2^-18 of the largest coordinate magnitude (covers the
    // test's plane shift of 3 |o| * 2^-24 for origins up to 20 times it), plus a tiny absolute term.
    float maxc = 0.0f;
    for (size_t i = 0; i < (size_t)num_tris * 9; i++) maxc = std::max(maxc, std::fabs(tri9[i]));
    const float pad = maxc * (1.0f / 262144.0f) + 1e-30f;
    Collapser c(b2.nodes_, pad);
    c.run();
    // Renumber the nodes in breadth-first order (the root stays 0), so the nodes nearest the root,
    // which most rays visit, form the prefix that the kernels keep in LDS.
    const std::vector<Bvh4Node>& dfs = c.out_;
    std::vector<uint32_t> bfs;
    bfs.reserve(dfs.size());
    bfs.push_back(0);
    for (size_t h = 0; h < bfs.size(); h++)
        for (int i = 0; i < 4; i++) {
            const uint32_t ch = dfs[bfs[h]].child[i];
            if (ch != RL_EMPTY_CHILD && !(ch & RL_LEAF_BIT)) bfs.push_back(ch);
        }
    std::vector<uint32_t> rank(dfs.size());
    for (size_t r = 0; r < bfs.size(); r++) rank[bfs[r]] = (uint32_t)r;
    out.nodes.resize(dfs.size());
    for (size_t r = 0; r < bfs.size(); r++) {
        Bvh4Node n = dfs[bfs[r]];
        for (int i = 0; i < 4; i++)
            if (n.child[i] != RL_EMPTY_CHILD && !(n.child[i] & RL_LEAF_BIT)) n.child[i] = rank[n.child[i]];
        out.nodes[r] = n;
    }
    out.order = std::move(b2.order_);
    out.stack_bound = c.stack_bound_;
    out.depth = c.depth_;
    return 0;
}

}  // namespace rl

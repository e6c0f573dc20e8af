// partition.cpp -- see partition.hpp for the reference semantics.
#include "partition.hpp"

#include <algorithm>
#include <cmath>
#include <memory>

namespace hbi {
namespace {

struct Node {
    int start = 0, size = 0;
    std::unique_ptr<Node> son[2];
    double bmin[3], bmax[3], diam = 0.0;
    bool leaf() const { return !son[0]; }
};

struct TreeBuilder {
    const double *crd[3];
    int minsz;
    std::vector<int> &lod;

    // bisection at the midpoint of the longest axis; the sweep keeps the
    // reference's swap order so the permutation is identical
    std::unique_ptr<Node> build(int start, int nd)
    {
        auto c = std::make_unique<Node>();
        c->start = start;
        c->size = nd;
        if (nd <= minsz) return c;
        int *l = lod.data() + start;
        double lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = hi[d] = crd[d][l[0]];
            for (int i = 1; i < nd; i++) {
                const double v = crd[d][l[i]];
                if (v < lo[d]) lo[d] = v;
                else if (hi[d] < v) hi[d] = v;
            }
        }
        int cut = 0;
        double span = hi[0] - lo[0];
        for (int d = 0; d < 3; d++)
            if (hi[d] - lo[d] > span) { span = hi[d] - lo[d]; cut = d; }
        const double mid = (hi[cut] + lo[cut]) / 2;
        const double *v = crd[cut];
        int a = 1, b = nd;   // 1-based cursors
        while (a < b) {
            while (a < nd && v[l[a - 1]] <= mid) a++;
            while (b >= 1 && v[l[b - 1]] > mid) b--;
            if (a < b) std::swap(l[a - 1], l[b - 1]);
        }
        c->son[0] = build(start, a - 1);
        c->son[1] = build(start + a - 1, nd - a + 1);
        return c;
    }

    void boxes(Node *c)
    {
        if (c->leaf()) {
            for (int d = 0; d < 3; d++) c->bmin[d] = c->bmax[d] = crd[d][lod[c->start]];
            for (int d = 0; d < 3; d++)
                for (int i = 1; i < c->size; i++) {
                    const double v = crd[d][lod[c->start + i]];
                    if (v < c->bmin[d]) c->bmin[d] = v;
                    if (c->bmax[d] < v) c->bmax[d] = v;
                }
        } else {
            boxes(c->son[0].get());
            boxes(c->son[1].get());
            for (int d = 0; d < 3; d++) { c->bmin[d] = c->son[0]->bmin[d]; c->bmax[d] = c->son[0]->bmax[d]; }
            for (int d = 0; d < 3; d++) {
                if (c->son[1]->bmin[d] < c->bmin[d]) c->bmin[d] = c->son[1]->bmin[d];
                if (c->bmax[d] < c->son[1]->bmax[d]) c->bmax[d] = c->son[1]->bmax[d];
            }
        }
        double w = (c->bmax[0] - c->bmin[0]) * (c->bmax[0] - c->bmin[0]);
        for (int d = 1; d < 3; d++) w = w + (c->bmax[d] - c->bmin[d]) * (c->bmax[d] - c->bmin[d]);
        c->diam = std::sqrt(w);
    }
};

struct Splitter {
    int leaf, nlmax;
    double eta;
    std::vector<Leaf> &out;

    void split(const Node *l, const Node *t, int depth)
    {
        depth += 1;
        const int nleaf = leaf + 1;
        double s = 0.0;
        for (int d = 0; d < 3; d++) {
            if (l->bmax[d] < t->bmin[d]) s = s + (t->bmin[d] - l->bmax[d]) * (t->bmin[d] - l->bmax[d]);
            else if (t->bmax[d] < l->bmin[d]) s = s + (l->bmin[d] - t->bmax[d]) * (l->bmin[d] - t->bmax[d]);
        }
        const double dist = std::sqrt(s);
        const int nr = l->size, nc = t->size;
        if ((l->diam <= eta * dist || t->diam <= eta * dist) && nr >= nleaf && nc >= nleaf && nr <= nlmax &&
            nc <= nlmax && (l->start + nr) != t->start && (t->start + nc) != l->start) {
            out.push_back({1, l->start, nr, t->start, nc, 0});
            return;
        }
        if (depth == 100 || l->leaf() || t->leaf() || nr <= nleaf || nc <= nleaf) {
            out.push_back({2, l->start, nr, t->start, nc, 0});
            return;
        }
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) split(l->son[a].get(), t->son[b].get(), depth);
    }
};

}  // namespace

Partition build_partition(int n, const double *x, const double *y, const double *z, int leaf, double eta)
{
    Partition p;
    p.lod.resize(n);
    for (int i = 0; i < n; i++) p.lod[i] = i;
    std::vector<double> zero;
    if (!y || !z) zero.assign(n, 0.0);
    TreeBuilder tb{{x, y ? y : zero.data(), z ? z : zero.data()}, leaf, p.lod};
    std::unique_ptr<Node> root = tb.build(0, n);
    tb.boxes(root.get());
    if (root->leaf() || n <= leaf + 1) {
        p.leaves.push_back({2, 0, n, 0, n, 0});
    } else {
        Splitter sp{leaf, n, eta, p.leaves};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) sp.split(root->son[a].get(), root->son[b].get(), 1);
    }
    std::sort(p.leaves.begin(), p.leaves.end(), [](const Leaf &u, const Leaf &v) {
        return u.r0 != v.r0 ? u.r0 < v.r0 : u.c0 < v.c0;
    });
    return p;
}

}  // namespace hbi

// bvh.cpp -- the reference's binary BVH (Src/bvh.cpp:63-314) built on the
// host and flattened in its DFS order with skip links
// (Src/meshmapping.cpp:62-125) into 32-byte device nodes.
//
// The tree is the reference's, node for node: 16 centroid buckets per axis,
// its surface-area cost (with the xy term counted twice and yz missing,
// Src/bvh.cpp:66), split at a bucket boundary, stable partition by centroid,
// leaf at <= max_tris or depth >= max_depth or when a side comes out empty,
// node boxes stored as centre/half-size and read back as centre -/+ size.
// Unlike the reference it keeps one index array partitioned in place (stable,
// through a scratch array) and builds disjoint subtrees on separate threads;
// the result does not depend on the thread count.
#include <algorithm>
#include <array>
#include <functional>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>

#include "host.h"

namespace fornos {
namespace {

struct Box {
    V3 minv = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    V3 maxv = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    void add(V3 p) { minv = hmin(p, minv); maxv = hmax(p, maxv); }  // BucketAABB::addPoint
};
// combine(aabb0, aabb1), Src/bvh.cpp:55-61
Box combine(const Box& a0, const Box& a1)
{
    Box r = a0;
    r.minv = hmin(a1.minv, r.minv);
    r.maxv = hmax(a1.maxv, r.maxv);
    return r;
}
// computeCost, Src/bvh.cpp:63-68
float cost(const Box& b, int count)
{
    V3 s = b.maxv - b.minv;
    float area = 2.0f * (s.x * s.y + s.x * s.z + s.x * s.y);
    return (float)count * area;
}

struct Node {
    V3 center, size;
    size_t subtree = 0;
    std::unique_ptr<Node[]> child;  // null or 2
    uint32_t begin = 0, end = 0;    // leaf triangle range in the index array
};

struct Builder {
    const fornos_mesh& m;
    size_t max_tris, max_depth;
    std::vector<V3> p0, p1, p2;     // triangle corners
    std::vector<uint32_t> idx, scratch;

    Builder(const fornos_mesh& mesh, size_t mt, size_t md) : m(mesh), max_tris(mt), max_depth(md)
    {
        const uint32_t nt = m.ntri();
        p0.resize(nt); p1.resize(nt); p2.resize(nt);
        for (uint32_t t = 0; t < nt; ++t) {
            p0[t] = m.pos_of(m.triangles[3 * t]);
            p1[t] = m.pos_of(m.triangles[3 * t + 1]);
            p2[t] = m.pos_of(m.triangles[3 * t + 2]);
        }
        idx.resize(nt);
        scratch.resize(nt);
        for (uint32_t t = 0; t < nt; ++t) idx[t] = t;
    }

    // selectSplitFromBuckets, Src/bvh.cpp:70-126.  At i == 15 the reference
    // reads one past aabbRightSide[15]; that cost term is multiplied by
    // countRight == 0 and contributes 0.
    static void select(const uint32_t bk[16], const Box bb[16], size_t count, int& best_idx, float& best_cost)
    {
        Box right[15];
        right[14] = bb[15];
        for (int i = 13; i >= 0; --i) right[i] = combine(right[i + 1], bb[i + 1]);
        best_cost = FLT_MAX;
        best_idx = 0;
        Box left;
        int cl = 0, cr = (int)count;
        for (int i = 0; i < 16; ++i) {
            cl += (int)bk[i];
            cr -= (int)bk[i];
            left = combine(left, bb[i]);
            float c = cost(left, cl) + (i < 15 ? cost(right[i], cr) : 0.0f);
            if (c < best_cost) { best_cost = c; best_idx = i; }
        }
    }

    // findBestSplit, Src/bvh.cpp:128-203
    void find_split(uint32_t b, uint32_t e, int& axis, float& split) const
    {
        Box cb;
        for (uint32_t k = b; k < e; ++k) {
            uint32_t t = idx[k];
            cb.add((p0[t] + p1[t] + p2[t]) / 3.0f);
        }
        uint32_t bx[16] = {0}, by[16] = {0}, bz[16] = {0};
        Box ax[16], ay[16], az[16];
        const V3 ext = cb.maxv - cb.minv;
        for (uint32_t k = b; k < e; ++k) {
            uint32_t t = idx[k];
            V3 c = (p0[t] + p1[t] + p2[t]) / 3.0f;
            V3 d = c - cb.minv;
            V3 ijk = mk(d.x / ext.x * 15.99f, d.y / ext.y * 15.99f, d.z / ext.z * 15.99f);
            size_t i = std::isnan(ijk.x) ? 0 : (size_t)ijk.x;
            size_t j = std::isnan(ijk.y) ? 0 : (size_t)ijk.y;
            size_t l = std::isnan(ijk.z) ? 0 : (size_t)ijk.z;
            ++bx[i]; ++by[j]; ++bz[l];
            ax[i].add(p0[t]); ax[i].add(p1[t]); ax[i].add(p2[t]);
            ay[j].add(p0[t]); ay[j].add(p1[t]); ay[j].add(p2[t]);
            az[l].add(p0[t]); az[l].add(p1[t]); az[l].add(p2[t]);
        }
        int ix, iy, iz;
        float cx, cy, cz;
        select(bx, ax, e - b, ix, cx);
        select(by, ay, e - b, iy, cy);
        select(bz, az, e - b, iz, cz);
        if (cx <= cy && cx <= cz) {
            axis = 0; split = cb.minv.x + (cb.maxv.x - cb.minv.x) / 16.0f * (float)(ix + 1);
        } else if (cy <= cx && cy <= cz) {
            axis = 1; split = cb.minv.y + (cb.maxv.y - cb.minv.y) / 16.0f * (float)(iy + 1);
        } else {
            axis = 2; split = cb.minv.z + (cb.maxv.z - cb.minv.z) / 16.0f * (float)(iz + 1);
        }
    }

    // binaryDivisionBVH, Src/bvh.cpp:205-281
    void divide(Node& par, size_t depth)
    {
        const uint32_t b = par.begin, e = par.end, n = e - b;
        if (n <= max_tris || depth >= max_depth) { par.subtree = n; return; }
        int axis;
        float split;
        find_split(b, e, axis, split);
        V3 minL = mk(FLT_MAX, FLT_MAX, FLT_MAX), maxL = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        V3 minR = minL, maxR = maxL;
        uint32_t nl = 0, nr = 0;
        for (uint32_t k = b; k < e; ++k) {
            uint32_t t = idx[k];
            V3 c = (p0[t] + p1[t] + p2[t]) * (1.0f / 3.0f);
            float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
            if (cv <= split) {
                idx[b + nl++] = t;  // stable: writes never overtake reads
                minL = hmin(minL, hmin(p0[t], hmin(p1[t], p2[t])));
                maxL = hmax(maxL, hmax(p0[t], hmax(p1[t], p2[t])));
            } else {
                scratch[b + nr++] = t;
                minR = hmin(minR, hmin(p0[t], hmin(p1[t], p2[t])));
                maxR = hmax(maxR, hmax(p0[t], hmax(p1[t], p2[t])));
            }
        }
        std::memcpy(&idx[b + nl], &scratch[b], sizeof(uint32_t) * nr);
        if (nl == 0 || nr == 0) { par.subtree = n; return; }  // "Unable to subdivide"
        par.child.reset(new Node[2]);
        Node& L = par.child[0];
        Node& R = par.child[1];
        L.begin = b; L.end = b + nl;
        R.begin = b + nl; R.end = e;
        L.center = (minL + maxL) * 0.5f; L.size = (maxL - minL) * 0.5f;
        R.center = (minR + maxR) * 0.5f; R.size = (maxR - minR) * 0.5f;
        if (n > 65536 && depth < 6) {
            auto f = std::async(std::launch::async, [&] { divide(L, depth + 1); });
            divide(R, depth + 1);
            f.get();
        } else {
            divide(L, depth + 1);
            divide(R, depth + 1);
        }
        V3 aminL = L.center - L.size, amaxL = L.center + L.size;
        V3 aminR = R.center - R.size, amaxR = R.center + R.size;
        V3 amin = hmin(aminL, aminR), amax = hmax(amaxL, amaxR);
        par.center = (amax + amin) * 0.5f;
        par.size = (amax - amin) * 0.5f;
        par.begin = par.end = 0;
        par.subtree = L.subtree + R.subtree;
    }
};

struct Flattener {
    const Builder& B;
    FlatBVH& out;
    uint32_t soup = 0;  // soup triangles emitted so far

    void put_f(size_t w, float f) { std::memcpy(&out.nodes[w], &f, 4); }

    // fillMeshData, Src/meshmapping.cpp:62-125
    void run(const Node& n, uint32_t depth)
    {
        const bool leaf = !n.child;
        if (leaf && n.end == n.begin) return;
        if (!leaf) {
            if (n.child[0].subtree > 0 && n.child[1].subtree == 0) { run(n.child[0], depth); return; }
            if (n.child[1].subtree > 0 && n.child[0].subtree == 0) { run(n.child[1], depth); return; }
        }
        const uint32_t index = out.n_nodes++;
        out.nodes.resize((size_t)out.n_nodes * 8);
        const V3 mn = n.center - n.size, mx = n.center + n.size;
        const size_t w = (size_t)index * 8;
        put_f(w + 0, mn.x); put_f(w + 1, mn.y); put_f(w + 2, mn.z);
        out.nodes[w + 3] = soup;
        put_f(w + 4, mx.x); put_f(w + 5, mx.y); put_f(w + 6, mx.z);
        out.depth = std::max(out.depth, depth);
        if (leaf) {
            const uint32_t cnt = n.end - n.begin;
            for (uint32_t k = n.begin; k < n.end; ++k) out.soup_tri[soup++] = B.idx[k];
            out.nodes[w + 7] = kLeafFlag | cnt;
            out.n_leaves++;
            out.max_leaf = std::max(out.max_leaf, cnt);
            return;
        }
        run(n.child[0], depth + 1);
        run(n.child[1], depth + 1);
        out.nodes[(size_t)index * 8 + 7] = out.n_nodes;
    }
};

}  // namespace

void build_reference_bvh(const fornos_mesh& m, uint32_t max_tris, uint32_t max_depth, FlatBVH& out)
{
    Builder B(m, max_tris, max_depth);
    // BVH::createBinary: root box over every mesh position (Src/bvh.cpp:288-300)
    V3 mins = mk(FLT_MAX, FLT_MAX, FLT_MAX), maxs = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (const V3& p : m.positions) { mins = hmin(mins, p); maxs = hmax(maxs, p); }
    Node root;
    root.center = (maxs + mins) * 0.5f;
    root.size = (maxs - mins) * 0.5f;
    root.begin = 0;
    root.end = m.ntri();
    B.divide(root, 0);
    out = FlatBVH();
    out.soup_tri.resize(m.ntri());
    out.nodes.reserve((size_t)16 * m.ntri() / std::max<uint32_t>(1, max_tris) + 16);
    Flattener F{B, out};
    F.run(root, 0);
    out.soup_tri.resize(F.soup);
}

void export_reference_nodes(const FlatBVH& B, uint32_t* nodes9)
{
    for (uint32_t i = 0; i < B.n_nodes; ++i) {
        const uint32_t* w = &B.nodes[8 * (size_t)i];
        uint32_t* o = &nodes9[9 * (size_t)i];
        o[0] = w[0]; o[1] = w[1]; o[2] = w[2]; o[3] = w[4]; o[4] = w[5]; o[5] = w[6];
        if (w[7] & kLeafFlag) {
            o[6] = 3 * w[3];
            o[7] = 3 * (w[3] + (w[7] & ~kLeafFlag));
            o[8] = i + 1;
        } else {
            o[6] = o[7] = 3 * w[3];
            o[8] = w[7];
        }
    }
}

std::vector<uint32_t> reference_parents(const FlatBVH& B)
{
    std::vector<uint32_t> parent(B.n_nodes, 0xFFFFFFFFu);
    for (uint32_t i = 0; i < B.n_nodes; ++i) {
        const uint32_t w = B.nodes[8 * (size_t)i + 7];
        if (w & kLeafFlag) continue;
        for (uint32_t c = i + 1; c < w;) {
            parent[c] = i;
            const uint32_t cw = B.nodes[8 * (size_t)c + 7];
            c = (cw & kLeafFlag) ? c + 1 : cw;
        }
    }
    return parent;
}

std::vector<uint32_t> reference_chain_links(const FlatBVH& B)
{
    const std::vector<uint32_t> parent = reference_parents(B);
    std::vector<uint32_t> link(B.n_nodes, 0xFFFFFFFFu);
    auto word = [&](uint32_t node, int k) {
        float f;
        std::memcpy(&f, &B.nodes[8 * (size_t)node + k], 4);
        return f;
    };
    // DFS order: a parent's index is below its children's
    for (uint32_t x = 0; x < B.n_nodes; ++x) {
        const uint32_t p = parent[x];
        if (p == 0xFFFFFFFFu) continue;
        bool nested = true;
        for (int k = 0; k < 3; ++k)
            nested = nested && word(p, k) <= word(x, k) && word(x, k) <= word(x, 4 + k) &&
                     word(x, 4 + k) <= word(p, 4 + k);
        link[x] = nested ? link[p] : p;
    }
    return link;
}

namespace {

float fbits(uint32_t w)
{
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

struct CullPrim {
    V3 lo, hi, c;
    uint32_t ref;       // reference leaf node
    uint32_t mask = 0;  // a group of its triangles (bit j = first + j), 0: all
};

float true_area(V3 lo, V3 hi)
{
    const V3 s = hi - lo;
    return 2.0f * (s.x * s.y + s.y * s.z + s.z * s.x);
}

struct CullNode {
    V3 lo, hi;
    int child = -1, right = -1;  // children in `pool`, or -1 for a leaf
    uint32_t prim = 0;         // leaf: primitive index
};

constexpr uint32_t kParallelPrims = 1u << 14;  // subtrees this large build their halves concurrently
constexpr uint32_t kParallelDepth = 4;         // down to this depth (at most 16 threads)

struct CullBuilder {
    std::vector<CullPrim>& prims;
    std::vector<CullNode> pool;

    // Binned SAH with the true surface area, 16 bins per axis, one primitive per leaf.
    int build(uint32_t b, uint32_t e, uint32_t depth, uint32_t& max_depth)
    {
        max_depth = std::max(max_depth, depth);
        const int id = (int)pool.size();
        pool.emplace_back();
        V3 lo = prims[b].lo, hi = prims[b].hi, clo = prims[b].c, chi = prims[b].c;
        for (uint32_t i = b + 1; i < e; ++i) {
            lo = hmin(lo, prims[i].lo); hi = hmax(hi, prims[i].hi);
            clo = hmin(clo, prims[i].c); chi = hmax(chi, prims[i].c);
        }
        pool[id].lo = lo;
        pool[id].hi = hi;
        if (e - b == 1) { pool[id].prim = b; return id; }
        const float ext[3] = {chi.x - clo.x, chi.y - clo.y, chi.z - clo.z};
        const float cl[3] = {clo.x, clo.y, clo.z};
        float best = FLT_MAX;
        int best_axis = -1, best_bin = 0;
        for (int ax = 0; ax < 3; ++ax) {
            if (!(ext[ax] > 0)) continue;
            uint32_t cnt[16] = {0};
            V3 blo[16], bhi[16];
            for (int k = 0; k < 16; ++k) { blo[k] = mk(FLT_MAX, FLT_MAX, FLT_MAX); bhi[k] = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX); }
            for (uint32_t i = b; i < e; ++i) {
                const float cv = ax == 0 ? prims[i].c.x : (ax == 1 ? prims[i].c.y : prims[i].c.z);
                int k = (int)((cv - cl[ax]) / ext[ax] * 16.0f);
                k = std::min(15, std::max(0, k));
                ++cnt[k];
                blo[k] = hmin(blo[k], prims[i].lo);
                bhi[k] = hmax(bhi[k], prims[i].hi);
            }
            float rarea[16];
            uint32_t rcnt[16];
            V3 rlo = mk(FLT_MAX, FLT_MAX, FLT_MAX), rhi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            uint32_t rc = 0;
            for (int k = 15; k >= 1; --k) {
                rlo = hmin(rlo, blo[k]); rhi = hmax(rhi, bhi[k]); rc += cnt[k];
                rarea[k] = rc ? true_area(rlo, rhi) : 0.0f;
                rcnt[k] = rc;
            }
            V3 llo = mk(FLT_MAX, FLT_MAX, FLT_MAX), lhi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            uint32_t lc = 0;
            for (int k = 0; k < 15; ++k) {
                llo = hmin(llo, blo[k]); lhi = hmax(lhi, bhi[k]); lc += cnt[k];
                if (lc == 0 || rcnt[k + 1] == 0) continue;
                const float c = true_area(llo, lhi) * (float)lc + rarea[k + 1] * (float)rcnt[k + 1];
                if (c < best) { best = c; best_axis = ax; best_bin = k; }
            }
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = b + (e - b) / 2;  // coincident centroids: split by index
        } else {
            auto it = std::partition(prims.begin() + b, prims.begin() + e, [&](const CullPrim& p) {
                const float cv = best_axis == 0 ? p.c.x : (best_axis == 1 ? p.c.y : p.c.z);
                int k = (int)((cv - cl[best_axis]) / ext[best_axis] * 16.0f);
                k = std::min(15, std::max(0, k));
                return k <= best_bin;
            });
            mid = (uint32_t)(it - prims.begin());
            if (mid == b || mid == e) mid = b + (e - b) / 2;
        }
        int l, r;
        if (e - b >= kParallelPrims && depth < kParallelDepth) {
            // the right half on another thread with a pool of its own (the
            // halves touch disjoint ranges of `prims`), grafted after ours:
            // the tree is the same as a sequential build's
            CullBuilder rb{prims, {}};
            uint32_t rdepth = 0;
            auto right = std::async(std::launch::async, [&] { rb.build(mid, e, depth + 1, rdepth); });
            l = build(b, mid, depth + 1, max_depth);
            right.get();
            r = (int)pool.size();
            for (CullNode n : rb.pool) {
                if (n.child >= 0) {
                    n.child += r;
                    n.right += r;
                }
                pool.push_back(n);
            }
            max_depth = std::max(max_depth, rdepth);
        } else {
            l = build(b, mid, depth + 1, max_depth);
            r = build(mid, e, depth + 1, max_depth);
        }
        pool[id].child = l;
        pool[id].right = r;
        return id;
    }
};

void emit_cull(const std::vector<CullNode>& pool, const std::vector<CullPrim>& prims, const FlatBVH& ref,
               int id, CullBVH& out)
{
    const CullNode& n = pool[id];
    const uint32_t index = out.n_nodes++;
    uint32_t w[8];
    std::memcpy(&w[0], &n.lo.x, 4); std::memcpy(&w[1], &n.lo.y, 4); std::memcpy(&w[2], &n.lo.z, 4);
    std::memcpy(&w[4], &n.hi.x, 4); std::memcpy(&w[5], &n.hi.y, 4); std::memcpy(&w[6], &n.hi.z, 4);
    w[3] = 0;
    w[7] = 0;
    if (n.child < 0) {
        const uint32_t* rw = &ref.nodes[8 * (size_t)prims[n.prim].ref];
        w[3] = rw[3];  // first soup triangle
        w[7] = rw[7];  // kLeafFlag | count
    }
    out.nodes.insert(out.nodes.end(), w, w + 8);
    out.ref_leaf.push_back(n.child < 0 ? prims[n.prim].ref : 0xFFFFFFFFu);
    if (n.child >= 0) {
        emit_cull(pool, prims, ref, n.child, out);
        emit_cull(pool, prims, ref, n.right, out);
        out.nodes[8 * (size_t)index + 7] = out.n_nodes;  // jump past the subtree
    }
}

}  // namespace

// Groups of at most 2 of a reference leaf's triangles (first, cnt), by
// median splits of their centroids on the widest axis, each with its padded
// box; appended to `parts`.
void triangle_groups(const std::vector<float>& spos4, uint32_t first, uint32_t cnt, float pad, uint32_t leaf,
                     std::vector<CullPrim>& parts)
{
    auto vtx = [&](uint32_t j, int c) {
        const size_t k = 3 * (size_t)(first + j) + c;
        return mk(spos4[4 * k], spos4[4 * k + 1], spos4[4 * k + 2]);
    };
    auto cen = [&](uint32_t j) { return (vtx(j, 0) + vtx(j, 1) + vtx(j, 2)) * (1.0f / 3.0f); };
    std::vector<uint32_t> idx(cnt);
    for (uint32_t j = 0; j < cnt; ++j) idx[j] = j;
    std::function<void(uint32_t, uint32_t)> split = [&](uint32_t b, uint32_t e) {
        if (e - b <= 2) {
            CullPrim cp;
            V3 l = mk(FLT_MAX, FLT_MAX, FLT_MAX), h = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            for (uint32_t q = b; q < e; ++q) {
                for (int c = 0; c < 3; ++c) {
                    l = hmin(l, vtx(idx[q], c));
                    h = hmax(h, vtx(idx[q], c));
                }
                cp.mask |= 1u << idx[q];
            }
            cp.lo = l - mk(pad, pad, pad);
            cp.hi = h + mk(pad, pad, pad);
            cp.c = (cp.lo + cp.hi) * 0.5f;
            cp.ref = leaf;
            parts.push_back(cp);
            return;
        }
        V3 cl = cen(idx[b]), ch = cl;
        for (uint32_t q = b + 1; q < e; ++q) {
            cl = hmin(cl, cen(idx[q]));
            ch = hmax(ch, cen(idx[q]));
        }
        const V3 ex = ch - cl;
        const int ax = (ex.x >= ex.y && ex.x >= ex.z) ? 0 : (ex.y >= ex.z ? 1 : 2);
        const uint32_t mid = b + (e - b) / 2;
        std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](uint32_t u, uint32_t v) {
            const V3 a = cen(u), c = cen(v);
            return ax == 0 ? a.x < c.x : (ax == 1 ? a.y < c.y : a.z < c.z);
        });
        split(b, mid);
        split(mid, e);
    };
    split(0, cnt);
}

void build_cull_bvh(const FlatBVH& ref, const std::vector<float>& spos4, float pad, CullBVH& out, float split_cost,
                    bool binary)
{
    // primitives of each reference leaf, leaves in node order; computed in
    // chunks of leaves on several threads and joined in the same order
    std::vector<uint32_t> leaves;
    leaves.reserve(ref.n_leaves);
    for (uint32_t i = 0; i < ref.n_nodes; ++i)
        if (ref.nodes[8 * (size_t)i + 7] & kLeafFlag) leaves.push_back(i);
    const size_t nchunk = std::max<size_t>(1, std::min<size_t>(64, leaves.size() / 1024));
    std::vector<std::vector<CullPrim>> chunk(nchunk);
    parallel_chunks(nchunk, [&](size_t c0, size_t c1) {
        std::vector<CullPrim> parts;
        for (size_t c = c0; c < c1; ++c) {
            const size_t l0 = leaves.size() * c / nchunk, l1 = leaves.size() * (c + 1) / nchunk;
            std::vector<CullPrim>& out_c = chunk[c];
            for (size_t li = l0; li < l1; ++li) {
                const uint32_t i = leaves[li];
                const uint32_t* w = &ref.nodes[8 * (size_t)i];
                V3 lo = mk(fbits(w[0]), fbits(w[1]), fbits(w[2])), hi = mk(fbits(w[4]), fbits(w[5]), fbits(w[6]));
                const uint32_t first = w[3], cnt = w[7] & ~kLeafFlag;
                if (split_cost > 0.0f && cnt >= 3 && cnt <= 32) {
                    // the groups replace the leaf when their expected triangle
                    // tests (each group's share of the leaf's surface area
                    // times its size) undercut the leaf's by more than split_cost
                    parts.clear();
                    triangle_groups(spos4, first, cnt, pad, i, parts);
                    V3 alo = parts[0].lo, ahi = parts[0].hi;
                    for (const CullPrim& cp : parts) {
                        alo = hmin(alo, cp.lo);
                        ahi = hmax(ahi, cp.hi);
                    }
                    const float area = true_area(alo, ahi);
                    float cost = split_cost;
                    for (const CullPrim& cp : parts)
                        cost += (area > 0 ? true_area(cp.lo, cp.hi) / area : 1.0f) * (float)__builtin_popcount(cp.mask);
                    if (cost < (float)cnt) {
                        out_c.insert(out_c.end(), parts.begin(), parts.end());
                        continue;
                    }
                }
                double edge = 0;  // longest triangle edge of the leaf
                for (uint32_t k = 3 * first; k < 3 * (first + cnt); ++k) {
                    const V3 p = mk(spos4[4 * k], spos4[4 * k + 1], spos4[4 * k + 2]);
                    lo = hmin(lo, p);
                    hi = hmax(hi, p);
                    const uint32_t k1 = k % 3 == 2 ? k - 2 : k + 1;
                    double e2 = 0;
                    for (int a = 0; a < 3; ++a) {
                        const double de = (double)spos4[4 * k + a] - spos4[4 * k1 + a];
                        e2 += de * de;
                    }
                    edge = std::max(edge, std::sqrt(e2));
                }
                // The mapping accepts barycentrics down to -1e-5
                // (Shaders/meshmapping.comp:75-81): its hits lie in the
                // triangle scaled by 1 + 3e-5 about its centroid, within
                // 3e-5 x the longest edge of the triangle.  With this margin
                // the box still holds them, so a ray enters it before any
                // such hit, and a traversal that prunes boxes entered at or
                // beyond its best hit keeps every leaf that could still win.
                const float m = pad + (float)(edge * (1.0 / 16384.0));
                CullPrim cp;
                cp.lo = lo - mk(m, m, m);
                cp.hi = hi + mk(m, m, m);
                cp.c = (cp.lo + cp.hi) * 0.5f;
                cp.ref = i;
                out_c.push_back(cp);
            }
        }
    }, 1);
    std::vector<CullPrim> prims;
    size_t total = 0;
    for (const auto& c : chunk) total += c.size();
    prims.reserve(total);
    for (auto& c : chunk) prims.insert(prims.end(), c.begin(), c.end());
    out = CullBVH();
    if (prims.empty()) return;
    CullBuilder cb{prims, {}};
    cb.pool.reserve(2 * prims.size());
    cb.build(0, (uint32_t)prims.size(), 0, out.depth);
    if (binary) {
        // DFS pre-order with skip links, in the node format of the reference layout
        out.nodes.reserve(8 * cb.pool.size());
        out.ref_leaf.reserve(cb.pool.size());
        emit_cull(cb.pool, prims, ref, 0, out);
    }
    // 4-wide collapse: open the largest-area internal child until 4 children
    struct Wide {
        const std::vector<CullNode>& pool;
        const std::vector<CullPrim>& prims;
        const FlatBVH& ref;
        CullBVH& out;
        uint32_t emit(int id)
        {
            std::vector<int> ch;
            const CullNode& n = pool[id];
            if (n.child < 0) ch.push_back(id);
            else { ch.push_back(n.child); ch.push_back(n.right); }
            while (ch.size() < 4) {
                int best = -1;
                float ba = -1;
                for (size_t i = 0; i < ch.size(); ++i) {
                    const CullNode& c = pool[ch[i]];
                    if (c.child < 0) continue;
                    const float a = true_area(c.lo, c.hi);
                    if (a > ba) { ba = a; best = (int)i; }
                }
                if (best < 0) break;
                const int open = ch[best];
                ch[best] = pool[open].child;
                ch.insert(ch.begin() + best + 1, pool[open].right);
            }
            const uint32_t index = out.n_wide++;
            out.wide.resize((size_t)32 * out.n_wide, 0);
            uint32_t w[32];
            for (int k = 0; k < 4; ++k) {
                float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
                uint32_t c = kNoChild;
                if (k < (int)ch.size()) {
                    const CullNode& cn = pool[ch[k]];
                    lo[0] = cn.lo.x; lo[1] = cn.lo.y; lo[2] = cn.lo.z;
                    hi[0] = cn.hi.x; hi[1] = cn.hi.y; hi[2] = cn.hi.z;
                    if (cn.child < 0) {
                        const uint32_t leaf = (uint32_t)(out.wide_leaves.size() / 4);
                        const uint32_t* rw = &ref.nodes[8 * (size_t)prims[cn.prim].ref];
                        const uint32_t lw[4] = {rw[3], rw[7] & ~kLeafFlag, prims[cn.prim].ref, prims[cn.prim].mask};
                        if (prims[cn.prim].mask) ++out.n_groups;
                        out.wide_leaves.insert(out.wide_leaves.end(), lw, lw + 4);
                        c = kLeafFlag | leaf;
                    } else {
                        c = emit(ch[k]);
                    }
                }
                for (int a = 0; a < 3; ++a) {
                    std::memcpy(&w[4 * a + k], &lo[a], 4);
                    std::memcpy(&w[12 + 4 * a + k], &hi[a], 4);
                }
                w[24 + k] = c;
                w[28 + k] = 0;
            }
            std::memcpy(&out.wide[(size_t)32 * index], w, sizeof(w));
            return index;
        }
    };
    Wide wd{cb.pool, prims, ref, out};
    wd.emit(0);
    // Surface-area cost of the wide tree: each node's box tests and each
    // leaf's triangle tests, weighted by the area of the box a ray has to
    // meet first, over the root's area.
    auto child_box = [&](const uint32_t* w, int k, V3& lo, V3& hi) {
        float v[6];
        for (int a = 0; a < 3; ++a) {
            std::memcpy(&v[a], &w[4 * a + k], 4);
            std::memcpy(&v[3 + a], &w[12 + 4 * a + k], 4);
        }
        lo = mk(v[0], v[1], v[2]);
        hi = mk(v[3], v[4], v[5]);
    };
    double root = 0, cost = 0;
    for (uint32_t n = 0; n < out.n_wide; ++n) {
        const uint32_t* w = &out.wide[32 * (size_t)n];
        V3 ulo = mk(FLT_MAX, FLT_MAX, FLT_MAX), uhi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        int nc = 0;
        for (int k = 0; k < 4; ++k) {
            if (w[24 + k] == kNoChild) continue;
            V3 lo, hi;
            child_box(w, k, lo, hi);
            ulo = hmin(ulo, lo);
            uhi = hmax(uhi, hi);
            ++nc;
            if (w[24 + k] & kLeafFlag) {
                const uint32_t* lf = &out.wide_leaves[4 * (size_t)(w[24 + k] & ~kLeafFlag)];
                cost += (double)true_area(lo, hi) * (lf[3] ? (uint32_t)__builtin_popcount(lf[3]) : lf[1]);
            }
        }
        const double a = true_area(ulo, uhi);
        if (n == 0) root = a;
        cost += a * kBoxTestCost * nc;
    }
    out.sah = root > 0 ? cost / root : 0;
    // Predicted cost were every whole leaf of 3..32 triangles split, without
    // building that tree: the leaf's triangle tests become one more level of
    // box tests plus its triangle pairs' (in soup order) share of them.
    double pred = cost;
    for (const CullPrim& cp : prims) {
        const uint32_t* rw = &ref.nodes[8 * (size_t)cp.ref];
        const uint32_t first = rw[3], cnt = rw[7] & ~kLeafFlag;
        if (cp.mask || cnt < 3 || cnt > 32) continue;
        const double al = true_area(cp.lo, cp.hi);
        double g = 0;
        uint32_t ng = 0;
        for (uint32_t j = 0; j < cnt; j += 2) {
            const uint32_t m = std::min(2u, cnt - j);
            V3 l = mk(FLT_MAX, FLT_MAX, FLT_MAX), h = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            for (uint32_t q = j; q < j + m; ++q)
                for (int c = 0; c < 3; ++c) {
                    const size_t k = 3 * (size_t)(first + q) + c;
                    const V3 p = mk(spos4[4 * k], spos4[4 * k + 1], spos4[4 * k + 2]);
                    l = hmin(l, p);
                    h = hmax(h, p);
                }
            g += (double)true_area(l - mk(pad, pad, pad), h + mk(pad, pad, pad)) * m;
            ++ng;
        }
        pred += g - al * cnt + al * kBoxTestCost * std::min(4u, ng);
    }
    out.sah_split_pred = root > 0 ? pred / root : 0;
}

// Group boxes inside the whole leaves of a culling tree: each wide leaf of
// 3..32 triangles gets a node of up to 4 groups (two rounds of centroid
// median splits), in the wide node format with a triangle mask in place of
// each child, when the groups' surface-area share of its triangle tests plus
// the four box tests (kBoxTestCost each) undercuts testing all of them.
void add_leaf_groups(CullBVH& t, const std::vector<float>& spos4, float pad)
{
    auto vtx = [&](uint32_t tri, int c) {
        const size_t k = 3 * (size_t)tri + c;
        return mk(spos4[4 * k], spos4[4 * k + 1], spos4[4 * k + 2]);
    };
    const size_t nleaf = t.wide_leaves.size() / 4;
    std::vector<std::array<uint32_t, 32>> node(nleaf);  // each leaf's group node (kept when word 24..27 != 0)
    std::vector<uint8_t> keep(nleaf, 0);
    parallel_chunks(nleaf, [&](size_t L0, size_t L1) {
        std::vector<uint32_t> idx;
        for (size_t L = L0; L < L1; ++L) {
            const uint32_t* lw = &t.wide_leaves[4 * L];
            const uint32_t first = lw[0], cnt = lw[1];
            if (cnt < 3 || cnt > 32) continue;
            auto cen = [&](uint32_t j) {
                return (vtx(first + j, 0) + vtx(first + j, 1) + vtx(first + j, 2)) * (1.0f / 3.0f);
            };
            idx.resize(cnt);
            for (uint32_t j = 0; j < cnt; ++j) idx[j] = j;
            auto halve = [&](uint32_t b, uint32_t e) {
                V3 cl = cen(idx[b]), ch = cl;
                for (uint32_t q = b + 1; q < e; ++q) {
                    cl = hmin(cl, cen(idx[q]));
                    ch = hmax(ch, cen(idx[q]));
                }
                const V3 ex = ch - cl;
                const int ax = (ex.x >= ex.y && ex.x >= ex.z) ? 0 : (ex.y >= ex.z ? 1 : 2);
                const uint32_t mid = b + (e - b) / 2;
                std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](uint32_t u, uint32_t v) {
                    const V3 a = cen(u), c = cen(v);
                    return ax == 0 ? a.x < c.x : (ax == 1 ? a.y < c.y : a.z < c.z);
                });
                return mid;
            };
            const uint32_t m = halve(0, cnt);
            const uint32_t cuts[5] = {0, m >= 2 ? halve(0, m) : m, m, cnt - m >= 2 ? halve(m, cnt) : cnt, cnt};
            V3 glo[4], ghi[4];
            uint32_t gmask[4] = {0, 0, 0, 0};
            V3 alo = mk(FLT_MAX, FLT_MAX, FLT_MAX), ahi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            for (int g = 0; g < 4; ++g) {
                glo[g] = mk(FLT_MAX, FLT_MAX, FLT_MAX);
                ghi[g] = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
                for (uint32_t q = cuts[g]; q < cuts[g + 1]; ++q) {
                    for (int c = 0; c < 3; ++c) {
                        glo[g] = hmin(glo[g], vtx(first + idx[q], c));
                        ghi[g] = hmax(ghi[g], vtx(first + idx[q], c));
                    }
                    gmask[g] |= 1u << idx[q];
                }
                if (!gmask[g]) continue;
                glo[g] = glo[g] - mk(pad, pad, pad);
                ghi[g] = ghi[g] + mk(pad, pad, pad);
                alo = hmin(alo, glo[g]);
                ahi = hmax(ahi, ghi[g]);
            }
            const float area = true_area(alo, ahi);
            float cost = 0;
            int ng = 0;
            for (int g = 0; g < 4; ++g)
                if (gmask[g]) {
                    cost += (area > 0 ? true_area(glo[g], ghi[g]) / area : 1.0f) * (float)__builtin_popcount(gmask[g]);
                    ++ng;
                }
            cost += (float)(kBoxTestCost * ng);
            if (!(cost < (float)cnt)) continue;
            uint32_t* w = node[L].data();
            for (int g = 0; g < 4; ++g) {
                float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
                if (gmask[g]) {
                    lo[0] = glo[g].x; lo[1] = glo[g].y; lo[2] = glo[g].z;
                    hi[0] = ghi[g].x; hi[1] = ghi[g].y; hi[2] = ghi[g].z;
                }
                for (int a = 0; a < 3; ++a) {
                    std::memcpy(&w[4 * a + g], &lo[a], 4);
                    std::memcpy(&w[12 + 4 * a + g], &hi[a], 4);
                }
                w[24 + g] = gmask[g];
                w[28 + g] = 0;
            }
            keep[L] = 1;
        }
    });
    // group nodes in leaf order
    t.sub.clear();
    t.n_sub = 0;
    for (size_t L = 0; L < nleaf; ++L) {
        t.wide_leaves[4 * L + 3] = 0;
        if (!keep[L]) continue;
        t.sub.insert(t.sub.end(), node[L].begin(), node[L].end());
        t.wide_leaves[4 * L + 3] = ++t.n_sub;
    }
}

}  // namespace fornos

extern "C" fornos_status fornos_bvh_reference_export(const fornos_mesh* m, uint32_t max_tris,
                                                     uint32_t max_depth, uint32_t* n_nodes,
                                                     uint32_t* nodes9, uint32_t* ids)
{
    if (!m || !n_nodes || max_tris == 0)
        return fornos::set_error(FORNOS_INVALID_ARG, "fornos_bvh_reference_export: bad argument");
    fornos::FlatBVH B;
    fornos::build_reference_bvh(*m, max_tris, max_depth, B);
    if (nodes9) {
        if (*n_nodes < B.n_nodes)
            return fornos::set_error(FORNOS_INVALID_ARG, "fornos_bvh_reference_export: node buffer too small");
        fornos::export_reference_nodes(B, nodes9);
        if (ids) std::memcpy(ids, B.soup_tri.data(), sizeof(uint32_t) * B.soup_tri.size());
    }
    *n_nodes = B.n_nodes;
    return FORNOS_OK;
}

extern "C" fornos_status fornos_bvh_reference_links(const fornos_mesh* m, uint32_t max_tris, uint32_t max_depth,
                                                    uint32_t n_nodes, uint32_t* parent, uint32_t* links)
{
    if (!m || max_tris == 0 || (!parent && !links))
        return fornos::set_error(FORNOS_INVALID_ARG, "fornos_bvh_reference_links: bad argument");
    fornos::FlatBVH B;
    fornos::build_reference_bvh(*m, max_tris, max_depth, B);
    if (n_nodes != B.n_nodes)
        return fornos::set_error(FORNOS_INVALID_ARG, "fornos_bvh_reference_links: node count differs from the tree's");
    if (parent) {
        const std::vector<uint32_t> p = fornos::reference_parents(B);
        std::memcpy(parent, p.data(), sizeof(uint32_t) * p.size());
    }
    if (links) {
        const std::vector<uint32_t> l = fornos::reference_chain_links(B);
        std::memcpy(links, l.data(), sizeof(uint32_t) * l.size());
    }
    return FORNOS_OK;
}

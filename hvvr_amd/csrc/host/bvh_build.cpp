// SAH bounding volume hierarchy construction for one mesh.
//
// Follows the three stages of shared/bvh.cpp (GenerateTopology, :542-556):
//   1. top-down binary SAH build over triangle boxes sorted by centroid
//      (scanAxis :167-194, helpBuildBVH2 :196-249, buildBVH2 :251-294);
//   2. post-order collapse of the binary tree into 4-wide nodes, evicting the
//      candidates whose removal gives the smallest box (helpBuildBVH4 :343-425,
//      buildBVH4 :427-467);
//   3. depth-first emission of TopologyNodes with triangles and vertices reordered
//      into first-use order (TopologyBuilder :474-540).
// Ties: the reference uses std::sort, whose order for equal keys is
// implementation-defined (MSVC introsort). Here every sort is stable, so equal keys
// keep their incoming order; for the <= 6 element sorts this matches MSVC's insertion
// sort exactly, for the large centroid sorts it is this library's definition.

#include "host_internal.h"

#include <algorithm>
#include <limits>

namespace hvvr {

namespace {

const float kInf = std::numeric_limits<float>::infinity();

struct Box {
    vector3 lower, upper;
    uint32_t offset = 0; // child node offset (0 = leaf)
    uint32_t size = 0;   // triangle count under this box
    static Box empty() {
        Box b;
        b.lower = vector3(kInf);
        b.upper = vector3(-kInf);
        return b;
    }
    void grow(const Box& o) {
        lower = vmin(lower, o.lower);
        upper = vmax(upper, o.upper);
    }
    float surfaceArea() const {
        vector3 e = upper - lower;
        return e.x * (e.y + e.z) + e.y * e.z;
    }
};

struct Node2 {
    Box child[2];
    Box bound() const {
        Box b = child[0];
        b.grow(child[1]);
        return b;
    }
};

float centroidKey(const Box* b, int axis) {
    return b->lower[axis] + b->upper[axis];
}

void sortAlong(const Box** boxes, size_t n, int axis) {
    std::stable_sort(boxes, boxes + n,
                     [axis](const Box* a, const Box* b) { return centroidKey(a, axis) < centroidKey(b, axis); });
}

// Evaluate all split positions along one axis; updates bestCost/bestIndex when a
// strictly cheaper split is found and reports whether it did.
bool evaluateAxis(const Box** boxes, size_t n, int axis, float* suffixCost, float& bestCost, size_t& bestIndex) {
    sortAlong(boxes, n, axis);
    Box acc = Box::empty();
    for (size_t i = 1; i < n; i++) {
        acc.grow(*boxes[n - i]);
        suffixCost[n - i] = acc.surfaceArea() * float(i);
    }
    bool improved = false;
    acc = Box::empty();
    for (size_t i = 1; i < n; i++) {
        acc.grow(*boxes[i - 1]);
        float cost = suffixCost[i] + acc.surfaceArea() * float(i);
        if (bestCost > cost) {
            bestCost = cost;
            bestIndex = i;
            improved = true;
        }
    }
    return improved;
}

Box unionRange(const Box* const* boxes, size_t n) {
    Box b = *boxes[0];
    for (size_t i = 1; i < n; i++)
        b.grow(*boxes[i]);
    return b;
}

// Returns the number of binary nodes written at `node` (0 = make this range a leaf).
size_t buildBinary(Node2* node, float* scratch, const Box** boxes, size_t n) {
    if (n < 2)
        return 0;
    int axis = 0;
    size_t split = 0;
    float cost = kInf;
    evaluateAxis(boxes, n, 0, scratch, cost, split);
    if (evaluateAxis(boxes, n, 1, scratch, cost, split))
        axis = 1;
    if (evaluateAxis(boxes, n, 2, scratch, cost, split))
        axis = 2;
    sortAlong(boxes, n, axis);

    size_t used = 1;
    size_t lowCount = buildBinary(node + used, scratch, boxes, split);
    if (lowCount) {
        node->child[0] = node[used].bound();
        node->child[0].offset = uint32_t(used);
        used += lowCount;
    } else {
        node->child[0] = unionRange(boxes, split);
        node->child[0].offset = 0;
    }
    node->child[0].size = uint32_t(split);

    size_t highCount = buildBinary(node + used, scratch + split, boxes + split, n - split);
    if (highCount) {
        node->child[1] = node[used].bound();
        node->child[1].offset = uint32_t(used);
        used += highCount;
    } else {
        node->child[1] = unionRange(boxes + split, n - split);
        node->child[1].offset = 0;
    }
    node->child[1].size = uint32_t(n - split);

    // a node costs as much as intersecting two triangles: keep the subdivision only if
    // it beats testing every triangle directly
    const float nodeCostInTriangles = 2.0f;
    return node->bound().surfaceArea() * (float(n) - nodeCostInTriangles) < cost ? 0 : used;
}

// candidate entry for the 4-wide collapse
struct Cand {
    Box box;
    uint32_t firstTri = 0, endTri = 0;
    uint32_t node = 0; // absolute index of the 4-wide node (0 = triangle range)
    uint32_t depth = 0;
};
struct Node4 {
    Cand child[4];
};

void sortChildren(Cand* c) {
    // leaves first, then by surface area, largest first (bvh.cpp:419-422)
    std::stable_sort(c, c + 4, [](const Cand& a, const Cand& b) {
        bool aLeaf = a.node == 0, bLeaf = b.node == 0;
        if (aLeaf != bLeaf)
            return aLeaf;
        return a.box.surfaceArea() > b.box.surfaceArea();
    });
}

struct Collapser {
    std::vector<Node4>& out;
    size_t next; // nodes are allocated from the back of `out`
    explicit Collapser(std::vector<Node4>& o) : out(o), next(o.size()) {}

    // Post-order: returns how many candidates were appended to `parent`.
    uint32_t collapse(const Node2* node, Cand* parent, uint32_t first) {
        Cand cache[6];
        uint32_t count = 0;
        uint32_t starts[2] = {first, first + node->child[0].size};
        for (int side = 0; side < 2; side++) {
            const Box& b = node->child[side];
            if (b.offset) {
                count += collapse(node + b.offset, cache + count, starts[side]);
            } else {
                Cand c;
                c.box = b;
                c.firstTri = starts[side];
                c.endTri = starts[side] + b.size;
                cache[count++] = c;
            }
        }
        if (count < 4) {
            for (uint32_t i = 0; i < count; i++)
                parent[i] = cache[i];
            return count;
        }
        uint32_t evictA = 4, evictB = 5;
        if (count > 4) {
            // prefer passing deeper candidates up, then keep the 4 with the smallest union
            std::stable_sort(cache, cache + count, [](const Cand& a, const Cand& b) { return a.depth > b.depth; });
            float best = kInf;
            if (count == 5) {
                for (uint32_t a = 0; a < 5; a++) {
                    Box u = Box::empty();
                    for (uint32_t i = 0; i < 5; i++)
                        if (i != a)
                            u.grow(cache[i].box);
                    if (best > u.surfaceArea()) {
                        best = u.surfaceArea();
                        evictA = a;
                    }
                }
            } else {
                for (uint32_t b = 1; b < 6; b++)
                    for (uint32_t a = 0; a < b; a++) {
                        Box u = Box::empty();
                        for (uint32_t i = 0; i < 6; i++)
                            if (i != a && i != b)
                                u.grow(cache[i].box);
                        if (best > u.surfaceArea()) {
                            best = u.surfaceArea();
                            evictA = a;
                            evictB = b;
                        }
                    }
            }
        }
        uint32_t self = uint32_t(--next);
        Node4& n4 = out[self];
        Cand* self_entry = parent++;
        Cand me;
        me.box = Box::empty();
        me.node = self;
        uint32_t slot = 0;
        for (uint32_t i = 0; i < count; i++) {
            if (i == evictA || i == evictB) {
                *parent++ = cache[i];
                continue;
            }
            if (cache[i].depth + 1 > me.depth)
                me.depth = cache[i].depth + 1;
            me.box.grow(cache[i].box);
            n4.child[slot++] = cache[i];
        }
        *self_entry = me;
        sortChildren(n4.child);
        return count - 3;
    }
};

struct TopologyWriter {
    const std::vector<Node4>& in;
    const PrecomputedTriangleShade* src;
    std::vector<PrecomputedTriangleShade> tris;
    std::vector<TopologyNode> nodes;
    TopologyWriter(const std::vector<Node4>& i, const PrecomputedTriangleShade* s) : in(i), src(s) {}

    uint32_t emit(uint32_t index) {
        uint32_t self = uint32_t(nodes.size());
        nodes.emplace_back();
        nodes[self].leafMask = 0;
        for (int i = 0; i < 7; i++)
            nodes[self].data[i] = 0;
        nodes[self].data[0] = uint32_t(tris.size());
        for (int k = 0; k < 4; k++) {
            const Cand& c = in[index].child[k];
            if (c.node) {
                uint32_t child = emit(c.node);
                nodes[self].data[1 + k] = child - self;
                continue;
            }
            nodes[self].leafMask |= 1u << k;
            for (uint32_t t = c.firstTri; t != c.endTri; t++)
                tris.push_back(src[t]);
            nodes[self].data[1 + k] = uint32_t(tris.size());
        }
        return self;
    }
};

} // namespace

bool GenerateTopology(MeshData& mesh, std::string* error) {
    size_t triCount = mesh.triShade.size();
    if (triCount == 0) {
        if (error)
            *error = "GenerateTopology: mesh has no triangles";
        return false;
    }
    for (const auto& t : mesh.triShade)
        for (int v = 0; v < 3; v++)
            if (t.indices[v] >= mesh.verts.size()) {
                if (error)
                    *error = "GenerateTopology: vertex index out of range";
                return false;
            }

    // 1. binary SAH build
    std::vector<Box> leafBoxes(triCount);
    std::vector<const Box*> order(triCount);
    for (size_t i = 0; i < triCount; i++) {
        const PrecomputedTriangleShade& t = mesh.triShade[i];
        vector3 p0 = mesh.verts[t.indices[0]].pos;
        vector3 p1 = mesh.verts[t.indices[1]].pos;
        vector3 p2 = mesh.verts[t.indices[2]].pos;
        leafBoxes[i].lower = vmin(p0, vmin(p1, p2));
        leafBoxes[i].upper = vmax(p0, vmax(p1, p2));
        order[i] = &leafBoxes[i];
    }
    std::vector<float> scratch(triCount);
    std::vector<Node2> node2s(triCount > 1 ? triCount - 1 : 1);
    size_t count = buildBinary(node2s.data(), scratch.data(), order.data(), triCount);
    if (!count) {
        Box all = unionRange(order.data(), triCount);
        node2s[0].child[0] = Box::empty();
        node2s[0].child[0].offset = 0;
        node2s[0].child[0].size = 0;
        node2s[0].child[1] = all;
        node2s[0].child[1].offset = 0;
        node2s[0].child[1].size = uint32_t(triCount);
        count = 1;
    }
    node2s.resize(count);
    std::vector<PrecomputedTriangleShade> sorted(triCount);
    for (size_t i = 0; i < triCount; i++)
        sorted[i] = mesh.triShade[size_t(order[i] - leafBoxes.data())];

    // 2. collapse to 4-wide
    std::vector<Node4> node4s((node2s.size() + 2) / 3);
    Collapser col(node4s);
    Cand top[3];
    uint32_t topCount = col.collapse(node2s.data(), top, 0);
    if (topCount != 1) {
        if (col.next != 1) {
            if (error)
                *error = "GenerateTopology: inconsistent 4-wide node count";
            return false;
        }
        Node4& root = node4s[--col.next];
        uint32_t slot = 0;
        for (uint32_t e = 4 - topCount; e; --e) {
            Cand empty; // zero box, empty triangle range
            empty.box.lower = vector3(0.0f);
            empty.box.upper = vector3(0.0f);
            root.child[slot++] = empty;
        }
        for (uint32_t i = 0; i < topCount; i++)
            root.child[slot++] = top[i];
        sortChildren(root.child);
    } else if (col.next != 0) {
        if (error)
            *error = "GenerateTopology: inconsistent 4-wide root";
        return false;
    }
    // child references become absolute indices; the writer resolves them

    // 3. depth-first topology, triangles in emission order
    TopologyWriter w(node4s, sorted.data());
    w.tris.reserve(triCount);
    w.emit(0);
    if (w.tris.size() != triCount || w.nodes.size() != node4s.size()) {
        if (error)
            *error = "GenerateTopology: topology does not cover every triangle";
        return false;
    }

    // vertices renumbered so the last triangle's last vertex is the last vertex
    uint32_t vertexCount = uint32_t(mesh.verts.size());
    std::vector<ShadingVertex> verts(vertexCount);
    std::vector<uint32_t> newIndex(vertexCount, vertexCount);
    uint32_t next = vertexCount;
    for (size_t t = w.tris.size(); t-- > 0;) {
        for (int v = 2; v >= 0; v--) {
            uint32_t old = w.tris[t].indices[v];
            if (newIndex[old] == vertexCount) {
                newIndex[old] = --next;
                verts[next] = mesh.verts[old];
            }
            w.tris[t].indices[v] = newIndex[old];
        }
    }
    if (next != 0) {
        if (error)
            *error = "GenerateTopology: mesh has vertices no triangle references";
        return false;
    }
    mesh.triShade = std::move(w.tris);
    mesh.verts = std::move(verts);
    mesh.nodes = std::move(w.nodes);
    return true;
}

bool GenerateTopology(MeshData& mesh) {
    std::string err;
    if (!GenerateTopology(mesh, &err))
        throw RaycasterError(err);
    return true;
}

} // namespace hvvr

// The host side of the centerline graph: runNewCenterlineAlg after the
// device stages (parallelCenterlineExtraction.cpp:1160-1316), loop removal
// and the centerline voxels (parallelCenterlineExtraction.cpp:27-322), the
// VTK writer (inputOutput.cpp:48-64); and the segmentation drivers
// (segmentation.cpp:5-307), each one device call.
//
// Loop removal is a sequential graph algorithm on a few thousand vertices,
// host code in the reference as here.  Its result depends on the order of
// every adjacency list and on the ties of the priority queue, so the steps
// below keep the reference's: nodes of degree two are fused into the edges
// between their neighbours in vertex order, a Prim spanning tree is grown
// from node 0 (then from the last unvisited node, while any is left) with a
// std::priority_queue ordered by edge length, and the fused nodes are put
// back along the tree edges.
#include <cmath>
#include <deque>
#include <fstream>
#include <queue>
#include <unordered_map>
#include <unordered_set>

#include "tsf_dropin.hpp"

namespace tsf {

namespace {

[[noreturn]] void fail_range(const char* msg) { throw Error(TSF_ERR_RANGE, msg); }

// SIPL int3::distance: the float length of the integer difference
float idist(const int3& a, const int3& b) {
    const int dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return std::sqrt((float)(dx * dx + dy * dy + dz * dz));
}

long long poskey(const int3& p, const int3& size) {
    return (long long)p.x + (long long)p.y * size.x + (long long)p.z * size.x * size.y;
}

struct GEdge;
struct GNode {
    int3 pos;
    std::vector<GEdge*> edges;
};
struct GEdge {
    GNode* source;
    GNode* target;
    std::vector<GNode*> removed;  // fused nodes from source to target
    float distance;
};
struct Longer {
    bool operator()(const GEdge* a, const GEdge* b) const { return a->distance > b->distance; }
};

// Owns every node and edge of one removeLoops call.
struct Arena {
    std::deque<GNode> nodes;
    std::deque<GEdge> edges;
    GEdge* edge(GNode* s, GNode* t, float d) {
        edges.push_back(GEdge{s, t, {}, d});
        return &edges.back();
    }
};

// drops every edge of n that ends at `target`, keeping the others' order
void drop_edges_to(GNode* n, const GNode* target) {
    std::vector<GEdge*> keep;
    for (GEdge* e : n->edges)
        if (e->target != target) keep.push_back(e);
    n->edges = keep;
}

// Prim from graph[root] over the nodes not yet visited; each tree edge ends
// up in its source's list, every other list of a reached node is emptied.
std::vector<GNode*> spanning_tree(int root, std::vector<GNode*>& graph, const int3& size,
                                 std::unordered_set<long long>& visited) {
    std::vector<GNode*> tree;
    std::priority_queue<GEdge*, std::vector<GEdge*>, Longer> queue;
    GNode* r = graph[root];
    tree.push_back(r);
    visited.insert(poskey(r->pos, size));
    for (GEdge* e : r->edges) queue.push(e);
    r->edges.clear();
    while (!queue.empty()) {
        GEdge* e = queue.top();
        queue.pop();
        if (visited.count(poskey(e->target->pos, size))) continue;
        for (GEdge* en : e->target->edges)
            if (!visited.count(poskey(en->target->pos, size))) queue.push(en);
        e->target->edges.clear();
        e->source->edges.push_back(e);
        tree.push_back(e->target);
        visited.insert(poskey(e->target->pos, size));
    }
    return tree;
}

// replaces tree edge e by the chain through its fused nodes
void unfuse(Arena& A, GEdge* e, std::vector<GNode*>& out) {
    if (e->removed.empty()) return;
    drop_edges_to(e->source, e->target);
    GNode* prev = e->source;
    for (GNode* k : e->removed) {
        prev->edges.push_back(A.edge(prev, k, 0.0f));
        prev = k;
        out.push_back(k);
    }
    prev->edges.push_back(A.edge(prev, e->target, 0.0f));
}

}  // namespace

void removeLoops(std::vector<int3>& vertices, std::vector<int2>& edges, int3& size) {
    Arena A;
    std::vector<GNode*> nodes;
    for (const int3& v : vertices) {
        A.nodes.push_back(GNode{v, {}});
        nodes.push_back(&A.nodes.back());
    }
    for (const int2& e : edges) {
        GNode* a = nodes[e.x];
        GNode* b = nodes[e.y];
        const float d = idist(a->pos, b->pos);
        a->edges.push_back(A.edge(a, b, d));
        b->edges.push_back(A.edge(b, a, d));
    }
    // fuse nodes of degree two into one edge between their neighbours
    std::vector<GNode*> graph;
    for (GNode* n : nodes) {
        if (n->edges.size() != 2) {
            graph.push_back(n);
            continue;
        }
        GEdge* e0 = n->edges[0];
        GEdge* e1 = n->edges[1];
        const float d = e0->distance + e1->distance;
        GNode* a = e0->target;
        GNode* b = e1->target;
        GEdge* ab = A.edge(a, b, d);
        ab->removed.assign(e0->removed.rbegin(), e0->removed.rend());
        ab->removed.push_back(n);
        ab->removed.insert(ab->removed.end(), e1->removed.begin(), e1->removed.end());
        GEdge* ba = A.edge(b, a, d);
        ba->removed.assign(e1->removed.rbegin(), e1->removed.rend());
        ba->removed.push_back(n);
        ba->removed.insert(ba->removed.end(), e0->removed.begin(), e0->removed.end());
        drop_edges_to(a, n);
        drop_edges_to(b, n);
        n->edges.clear();
        a->edges.push_back(ab);
        b->edges.push_back(ba);
    }
    if (graph.empty()) fail_range("Centerline graph size is 0! Can't continue. Maybe lower min-tree-length?");
    std::unordered_set<long long> visited;
    std::vector<GNode*> tree = spanning_tree(0, graph, size, visited);
    while (tree.size() < graph.size()) {
        int root = 0;
        for (int i = 0; i < (int)graph.size(); i++)
            if (!visited.count(poskey(graph[i]->pos, size))) root = i;
        std::vector<GNode*> more = spanning_tree(root, graph, size, visited);
        tree.insert(tree.end(), more.begin(), more.end());
    }
    // put the fused nodes back along the tree edges
    std::vector<GNode*> final_nodes;
    for (GNode* n : tree) {
        final_nodes.push_back(n);
        const std::vector<GEdge*> ne = n->edges;
        for (GEdge* e : ne) unfuse(A, e, final_nodes);
    }
    // renumber: nodes by first appearance (as a node or an edge target)
    std::vector<int3> nv;
    std::vector<int2> nedges;
    std::unordered_map<long long, int> index;
    auto id = [&](const GNode* n) {
        const long long k = poskey(n->pos, size);
        auto it = index.find(k);
        if (it != index.end()) return it->second;
        nv.push_back(n->pos);
        index[k] = (int)nv.size() - 1;
        return (int)nv.size() - 1;
    };
    for (GNode* n : final_nodes) {
        const int a = id(n);
        for (GEdge* e : n->edges) nedges.push_back(int2(a, id(e->target)));
    }
    vertices = nv;
    edges = nedges;
}

char* createCenterlineVoxels(std::vector<int3>& vertices, std::vector<int2>& edges, float* radius, int3& size) {
    const size_t total = (size_t)size.x * size.y * size.z;
    char* centerlines = new char[total]();
    auto pos = [&](const int3& a, const float (&dir)[3], float ratio) {
        const int x = (int)std::round((float)a.x + dir[0] * ratio), y = (int)std::round((float)a.y + dir[1] * ratio);
        const int z = (int)std::round((float)a.z + dir[2] * ratio);
        return (size_t)x + (size_t)size.x * ((size_t)y + (size_t)size.y * z);
    };
    for (const int2& e : edges) {
        const int3 a = vertices[e.x], b = vertices[e.y];
        const float distance = idist(a, b);
        const float dir[3] = {(float)(b.x - a.x), (float)(b.y - a.y), (float)(b.z - a.z)};
        const int n = (int)std::ceil(distance);
        float avg = 0.0f;
        for (int j = 0; j < n; j++) {
            const size_t q = pos(a, dir, (float)j / n);
            centerlines[q] = 1;
            avg += radius[q];
        }
        avg /= n;
        for (int j = 0; j < n; j++) radius[pos(a, dir, (float)j / n)] = avg;
    }
    return centerlines;
}

void writeToVtkFile(paramList& parameters, std::vector<int3> vertices, std::vector<int2> edges) {
    std::ofstream file(getParamStr(parameters, "centerline-vtk-file").c_str());
    file << "# vtk DataFile Version 3.0\nvtk output\nASCII\n";
    file << "DATASET POLYDATA\nPOINTS " << vertices.size() << " int\n";
    for (const int3& v : vertices) file << v.x << " " << v.y << " " << v.z << "\n";
    file << "\nLINES " << edges.size() << " " << edges.size() * 3 << "\n";
    for (const int2& e : edges) file << "2 " << e.x << " " << e.y << "\n";
}

Buffer runNewCenterlineAlg(OpenCL& ocl, int3 size, paramList& parameters, Image3D& vectorField, Image3D& TDF,
                           Image3D& radius) {
    CenterpointVertices V = runNewCenterlineAlgCandidates(ocl, size, parameters, vectorField, TDF, radius);
    if (getParamBool(parameters, "centerpoints-only")) return V.centerpoints;
    const tsf_dims d{size.x, size.y, size.z};
    const size_t total = (size_t)size.x * size.y * size.z;
    const int sum = V.count;
    const int minTreeLength = (int)getParam(parameters, "min-tree-length");
    // linking, duplicate removal and the edge list (parallelCenterlineExtraction.cpp:912-1060)
    Buffer lws(tsf_link_centerpoints_workspace_size(sum));
    int32_t sum2 = 0;
    check(tsf_link_centerpoints((const int32_t*)V.vertices.get(), sum, (const float*)TDF.plane(0), d,
                                getParam(parameters, "max-distance"), getParam(parameters, "min-mean-tdf"), nullptr,
                                0, &sum2, lws.get(), lws.size(), ocl.queue));
    if (sum2 == 0) fail_range("No edges were found");
    if (sum2 > 10000000) fail_range("More than 10 million edges found. Must be wrong!");
    Buffer edges((size_t)sum2 * 8);
    check(tsf_link_centerpoints((const int32_t*)V.vertices.get(), sum, (const float*)TDF.plane(0), d,
                                getParam(parameters, "max-distance"), getParam(parameters, "min-mean-tdf"),
                                (int32_t*)edges.get(), sum2, &sum2, lws.get(), lws.size(), ocl.queue));
    // components and their sizes (1062-1160)
    Buffer C((size_t)sum * 4), S((size_t)sum * 4);
    check(tsf_graph_components((const int32_t*)edges.get(), sum2, sum, (int32_t*)C.get(), (int32_t*)S.get(),
                               ocl.queue));
    Buffer centerlines(total);
    const bool vtk = getParamStr(parameters, "centerline-vtk-file") != "off";
    if (!vtk && !getParamBool(parameters, "loop-removal")) {
        // the device rasterisation of the large trees (1270-1316)
        check(tsf_memset(centerlines.get(), 0, total, ocl.queue));
        check(tsf_rasterize_tree_edges((const int32_t*)edges.get(), sum2, (const int32_t*)V.vertices.get(),
                                       (const int32_t*)C.get(), (const int32_t*)S.get(), minTreeLength,
                                       (uint8_t*)centerlines.get(), d, ocl.queue));
        return centerlines;
    }
    // on the host (1170-1268): the large trees' vertices and edges (long
    // edges split for the VTK file), loop removal, the centerline voxels
    std::vector<int> va((size_t)sum * 3), ea((size_t)sum2 * 2), Ca(sum), Sa(sum);
    V.vertices.download(ocl, va.data(), va.size() * 4);
    edges.download(ocl, ea.data(), ea.size() * 4);
    C.download(ocl, Ca.data(), Ca.size() * 4);
    S.download(ocl, Sa.data(), Sa.size() * 4);
    std::vector<float> radiusB = radius.downloadPlane(ocl, 0);
    std::vector<int3> vertices;
    std::vector<int> indexes(sum, -1);
    int counter = 0;
    for (int i = 0; i < sum; i++) {
        if (Sa[Ca[i]] >= minTreeLength) {
            vertices.push_back(int3(va[3 * i], va[3 * i + 1], va[3 * i + 2]));
            indexes[i] = counter++;
        }
    }
    std::vector<int2> graphEdges;
    const int maxEdgeDistance = (int)getParam(parameters, "max-edge-distance");
    for (int i = 0; i < sum2; i++) {
        const int ea0 = ea[2 * i], eb0 = ea[2 * i + 1];
        if (Sa[Ca[ea0]] < minTreeLength || Sa[Ca[eb0]] < minTreeLength) continue;
        const int3 A = vertices[indexes[ea0]], B = vertices[indexes[eb0]];
        const float distance = idist(A, B);
        if (vtk && distance > maxEdgeDistance) {
            const float dir[3] = {(float)(B.x - A.x), (float)(B.y - A.y), (float)(B.z - A.z)};
            int previous = indexes[ea0];
            for (int j = maxEdgeDistance; j < distance; j += maxEdgeDistance) {
                const float t = (float)j / distance;
                vertices.push_back(int3((int)std::round((float)A.x + t * dir[0]), (int)std::round((float)A.y + t * dir[1]),
                                        (int)std::round((float)A.z + t * dir[2])));
                graphEdges.push_back(int2(previous, counter));
                previous = counter++;
            }
            graphEdges.push_back(int2(previous, indexes[eb0]));
        } else {
            graphEdges.push_back(int2(indexes[ea0], indexes[eb0]));
        }
    }
    removeLoops(vertices, graphEdges, size);
    std::unique_ptr<char[]> cl(createCenterlineVoxels(vertices, graphEdges, radiusB.data(), size));
    centerlines.upload(ocl, cl.get(), total);
    radius.uploadPlane(ocl, 0, radiusB.data());
    if (vtk) writeToVtkFile(parameters, vertices, graphEdges);
    ocl.finish();
    return centerlines;
}

Buffer runInverseGradientSegmentation(OpenCL& ocl, Buffer& centerline, Image3D& vectorField, Image3D& radius,
                                      int3 size, paramList parameters) {
    (void)parameters;
    const tsf_dims d{size.x, size.y, size.z};
    const size_t total = (size_t)size.x * size.y * size.z;
    Buffer seg(total), ws(tsf_segment_inverse_gradient_workspace_size(d));
    tsf_field f = vectorField.field();
    int32_t passes = 0;
    check(tsf_segment_inverse_gradient((const uint8_t*)centerline.get(), (const float*)radius.plane(0), &f, d,
                                       (uint8_t*)seg.get(), &passes, ws.get(), ws.size(), ocl.queue));
    ocl.finish();
    return seg;
}

Buffer runSphereSegmentation(OpenCL& ocl, Buffer& centerline, Image3D& radius, int3 size, paramList parameters) {
    (void)parameters;
    const tsf_dims d{size.x, size.y, size.z};
    Buffer seg((size_t)size.x * size.y * size.z);
    check(tsf_segment_spheres((const uint8_t*)centerline.get(), (const float*)radius.plane(0), d,
                              (uint8_t*)seg.get(), ocl.queue));
    return seg;
}

}  // namespace tsf

// elas_tri_unit.cpp -- one Triangle instance, compiled with -DTRILIB=elas_triK.
//
// Includes the reference's src/triangle.cpp unmodified (found through -I) and
// adds one driver to its namespace.  elas_triangulate_arranged() takes the
// steps of triangulate() for the switches "zQB" (src/triangle.cpp:8503-8640:
// triangleinit, parsecommandline, transfernodes, delaunay, then the element
// list and triangledeinit) and calls Triangle's own divide-and-conquer
// (divconqrecurse, removeghosts).  The one step it does not take from
// Triangle is the vertex order that divconqdelaunay() builds with
// vertexsort() and alternateaxes() (src/triangle.cpp:6164-6230); the caller
// passes that order in, or (elas_triangulate_dedup) Triangle's own vertexsort()
// and duplicate removal run first and only the alternating cuts are built
// outside.  For distinct points the order is unique: both
// routines compare by (axis, other axis), a strict total order, each cut
// keeps the `divider` smallest points on its left, and every 2- or 3-point
// leaf ends sorted by (x, y).  Their random pivots only change how fast the
// order is reached, and the divide-and-conquer path draws no other random
// numbers, so the mesh and its triangle numbering are those of triangulate().
#include "triangle.cpp"

#include <stdint.h>

#include <vector>

#include "elas_triangle.h"

namespace TRILIB {

namespace {

// triangulate()'s steps up to the divide-and-conquer (src/triangle.cpp:
// 8503-8530 and delaunay(), :6237-6250); byinput[i] is input point i's vertex.
vertex* elas_begin(struct mesh& m, struct behavior& b, const float* pts, int n) {
    char sw[] = "zQB";
    char* swp = sw;
    triangleinit(&m);
    parsecommandline(1, &swp, &b);
    m.steinerleft = b.steiner;
    transfernodes(&m, &b, (float*)pts, (float*)NULL, (int*)NULL, n, 0);
    m.eextras = 0;
    initializetrisubpools(&m, &b);
    vertex* byinput = (vertex*)trimalloc(n * (int)sizeof(vertex));
    traversalinit(&m.vertices);
    for (int i = 0; i < n; i++) byinput[i] = vertextraverse(&m);
    return byinput;
}

// The merge on the arranged vertices (divconqdelaunay(), :6222-6230), then
// the element list.  triangulate() numbers the vertices in writenodes();
// without the point list numbernodes() gives the same numbers (no -j: every
// vertex counts).
int elas_finish(struct mesh& m, struct behavior& b, vertex* sortarray, int count, int** trilist) {
    struct otri hullleft, hullright;
    divconqrecurse(&m, &b, sortarray, count, 0, &hullleft, &hullright);
    trifree((int*)sortarray);
    const long hull = removeghosts(&m, &b, &hullleft);
    m.hullsize = m.triangles.items == 0 ? 0l : hull;
    numbernodes(&m, &b);
    float* attr = (float*)NULL;
    *trilist = (int*)NULL;
    writeelements(&m, &b, trilist, &attr);
    const int nt = (int)m.triangles.items;
    triangledeinit(&m, &b);
    return nt;
}

}  // namespace

// pts: [n][2], n >= 3, distinct; arr: the divide-and-conquer vertex order as
// input indices.  *trilist receives a malloc'ed [ntri][3] zero-based list
// (free() it); returns ntri.
int elas_triangulate_arranged(const float* pts, int n, const int* arr, int** trilist) {
    struct mesh m;
    struct behavior b;
    vertex* byinput = elas_begin(m, b, pts, n);
    vertex* sortarray = (vertex*)trimalloc(n * (int)sizeof(vertex));
    for (int k = 0; k < n; k++) sortarray[k] = byinput[arr[k]];
    trifree((int*)byinput);
    return elas_finish(m, b, sortarray, n, trilist);
}

// Any point set (n >= 3): Triangle's own vertexsort() on the caller's order
// and its duplicate removal (:6176-6198), which decide which copy of a
// repeated point survives; then the alternating-cut order of the survivors
// from `arrange`, which gets their coordinates in (x, y) order and returns
// positions into that list.
namespace {

// elas_begin(), Triangle's sort and duplicate removal, then the cuts of the
// survivors; returns the arranged vertices and their count, or NULL (mesh
// released) when a single distinct point is left.
vertex* elas_dedup_arrange(struct mesh& mesh, struct behavior& b, const float* pts, int n,
                           bool (*arrange)(const float* xy, int count, int* arr), int* count_out) {
    struct mesh* m = &mesh;  // Triangle's vertex macros name the mesh `m`
    vertex* sorted = elas_begin(mesh, b, pts, n);
    vertexsort(sorted, n);
    int i = 0;
    for (int j = 1; j < n; j++) {
        if (sorted[i][0] == sorted[j][0] && sorted[i][1] == sorted[j][1]) {
            setvertextype(sorted[j], UNDEADVERTEX);
            m->undeads++;
        } else {
            i++;
            sorted[i] = sorted[j];
        }
    }
    const int count = i + 1;
    if (count < 2) {
        // every point is the same point: Triangle's divconqrecurse() has no
        // case for a single vertex and recurses until the stack overflows;
        // report it as too few points instead (two distinct points give an
        // empty mesh, as in Triangle)
        trifree((int*)sorted);
        triangledeinit(&mesh, &b);
        return (vertex*)NULL;
    }
    float* xy = (float*)trimalloc(2 * count * (int)sizeof(float));
    int* arr = (int*)trimalloc(count * (int)sizeof(int));
    for (int k = 0; k < count; k++) {
        xy[2 * k] = sorted[k][0];
        xy[2 * k + 1] = sorted[k][1];
    }
    arrange(xy, count, arr);
    vertex* sortarray = (vertex*)trimalloc(count * (int)sizeof(vertex));
    for (int k = 0; k < count; k++) sortarray[k] = sorted[arr[k]];
    trifree((int*)xy);
    trifree(arr);
    trifree((int*)sorted);
    *count_out = count;
    return sortarray;
}

}  // namespace

int elas_triangulate_dedup(const float* pts, int n, bool (*arrange)(const float* xy, int count, int* arr),
                           int** trilist) {
    struct mesh mesh;
    struct behavior b;
    int count = 0;
    vertex* sortarray = elas_dedup_arrange(mesh, b, pts, n, arrange, &count);
    *trilist = (int*)NULL;
    if (!sortarray) return -1;
    return elas_finish(mesh, b, sortarray, count, trilist);
}


// ---- the divide-and-conquer split over threads ------------------------------
//
// divconqrecurse() on n vertices triangulates the first n/2 of the order,
// then the rest, then merges the two (src/triangle.cpp:6092-6104), so the
// triangles are allocated in post-order: left subtree, right subtree, the
// merge's own.  Only removeghosts() frees triangles, after the whole merge.
// The top `depth` levels of that recursion are cut out here: every node of
// the cut gets its own triangle pool (sharing the mesh's "outer space"
// triangle), the leaves run divconqrecurse() concurrently, the internal nodes
// run mergehulls() once both children are done, and the triangles are read
// out pool by pool in post-order, each pool in allocation order, skipping
// the dead ghosts -- the order in which writeelements() meets them in the
// single pool of a serial run.  Vertex handles, orientations and the merge
// arithmetic are Triangle's own, so the triangle list is that of
// triangulate().  Triangle's namespace-scope state (the exact-arithmetic
// bounds, the RNG) is written only by triangleinit() on the calling thread,
// before any task starts, and read-only during the merge.

namespace {

struct SplitNode {
    int lo, size, axis;
    int left, right;  // child nodes, -1 for a leaf
    struct mesh sub;  // the node's own triangle pool (the root uses the caller's mesh)
    struct otri farleft, farright;
};

struct SplitRun {
    struct behavior* b;
    vertex* sortarray;
    std::vector<SplitNode>* nodes;
    std::vector<int> leaves;
};

// post-order build of the cut: nodes of more than 3 vertices above `depth`
// split as divconqrecurse() splits them
int split_build(std::vector<SplitNode>& nodes, int lo, int size, int axis, int depth) {
    SplitNode nd;
    nd.lo = lo;
    nd.size = size;
    nd.axis = axis;
    nd.left = nd.right = -1;
    if (depth > 0 && size > 3) {
        const int divider = size >> 1;
        nd.left = split_build(nodes, lo, divider, 1 - axis, depth - 1);
        nd.right = split_build(nodes, lo + divider, size - divider, 1 - axis, depth - 1);
    }
    nodes.push_back(nd);
    return (int)nodes.size() - 1;
}

void split_leaf(void* arg, int i) {
    SplitRun* r = (SplitRun*)arg;
    SplitNode& nd = (*r->nodes)[r->leaves[i]];
    divconqrecurse(&nd.sub, r->b, r->sortarray + nd.lo, nd.size, nd.axis, &nd.farleft, &nd.farright);
}

// The merge on the arranged vertices cut into at most 2^depth subtrees, then
// the element list into *out ([ntri][3]); returns ntri.  depth 0 is the serial
// divconqrecurse().
int elas_finish_split(struct mesh& mesh, struct behavior& b, vertex* sortarray, int count, int depth,
                      ElasParallelFor par, void* parctx, std::vector<int32_t>* out) {
    struct mesh* m = &mesh;  // Triangle's vertex macros name the mesh `m`
    std::vector<SplitNode> nodes;
    nodes.reserve(2u << std::min(depth, 12));
    const int root = split_build(nodes, 0, count, 0, depth);
    SplitRun run;
    run.b = &b;
    run.sortarray = sortarray;
    run.nodes = &nodes;
    for (int i = 0; i < (int)nodes.size(); i++) {
        SplitNode& nd = nodes[i];
        if (i == root) continue;
        nd.sub = mesh;  // dummytri, sizes and flags of the mesh; a pool of its own:
        // room for the node's triangles (about 2 per vertex) in the first block
        poolinit(&nd.sub.triangles, mesh.triangles.itembytes, TRIPERBLOCK, 2 * nd.size + 8,
                 mesh.triangles.alignbytes);
        if (nd.left < 0) run.leaves.push_back(i);
    }
    if (nodes[root].left < 0) {
        nodes[root].sub = mesh;
        divconqrecurse(m, &b, sortarray, count, 0, &nodes[root].farleft, &nodes[root].farright);
    } else {
        par(parctx, (int)run.leaves.size(), split_leaf, &run);
        for (int i = 0; i < (int)nodes.size(); i++) {  // post-order: children before parents
            SplitNode& nd = nodes[i];
            if (nd.left < 0) continue;
            struct otri farleft = nodes[nd.left].farleft, innerleft = nodes[nd.left].farright;
            struct otri innerright = nodes[nd.right].farleft, farright = nodes[nd.right].farright;
            mergehulls(i == root ? m : &nd.sub, &b, &farleft, &innerleft, &innerright, &farright, nd.axis);
            nd.farleft = farleft;
            nd.farright = farright;
        }
    }
    trifree((int*)sortarray);
    removeghosts(m, &b, &nodes[root].farleft);
    numbernodes(m, &b);
    out->clear();
    out->reserve(6 * (size_t)count + 6);
    for (int i = 0; i < (int)nodes.size(); i++) {
        struct memorypool* pool = i == root ? &mesh.triangles : &nodes[i].sub.triangles;
        traversalinit(pool);
        struct otri t;
        t.orient = 0;
        vertex p1, p2, p3;
        while ((t.tri = (triangle*)traverse(pool)) != (triangle*)NULL) {
            if (deadtri(t.tri)) continue;
            org(t, p1);
            dest(t, p2);
            apex(t, p3);
            out->push_back(vertexmark(p1));
            out->push_back(vertexmark(p2));
            out->push_back(vertexmark(p3));
        }
        if (i != root) pooldeinit(pool);
    }
    triangledeinit(&mesh, &b);
    return (int)(out->size() / 3);
}

}  // namespace

// elas_triangulate_arranged() with the merge cut into at most 2^depth
// concurrent subtrees run through `par`; the triangles go to *out.
int elas_triangulate_split(const float* pts, int n, const int* arr, int depth, ElasParallelFor par, void* parctx,
                           std::vector<int32_t>* out) {
    struct mesh m;
    struct behavior b;
    vertex* byinput = elas_begin(m, b, pts, n);
    vertex* sortarray = (vertex*)trimalloc(n * (int)sizeof(vertex));
    for (int k = 0; k < n; k++) sortarray[k] = byinput[arr[k]];
    trifree((int*)byinput);
    return elas_finish_split(m, b, sortarray, n, depth, par, parctx, out);
}

// elas_triangulate_dedup() with the split merge; -1 for a single distinct point.
int elas_triangulate_dedup_split(const float* pts, int n, bool (*arrange)(const float* xy, int count, int* arr),
                                 int depth, ElasParallelFor par, void* parctx, std::vector<int32_t>* out) {
    struct mesh mesh;
    struct behavior b;
    int count = 0;
    vertex* sortarray = elas_dedup_arrange(mesh, b, pts, n, arrange, &count);
    out->clear();
    if (!sortarray) return -1;
    return elas_finish_split(mesh, b, sortarray, count, depth, par, parctx, out);
}

}  // namespace TRILIB

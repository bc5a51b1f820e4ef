// elas_tri_unit.cpp -- one Triangle instance, compiled with -DTRILIB=elas_triK.
//
// Includes the reference's src/triangle.cpp unmodified (found through -I) and
// adds one driver to its namespace (elas_tri_setup, elas_tri_finish).  It takes the
// steps of triangulate() for the switches "zQB" (src/triangle.cpp:8503-8640:
// triangleinit, parsecommandline, transfernodes, delaunay, then the element
// list and triangledeinit) and runs Triangle's own divide-and-conquer
// (divconqrecurse, mergehulls, removeghosts).  Two steps are its own:
//
//  * The vertex order that divconqdelaunay() builds with vertexsort() and
//    alternateaxes() (src/triangle.cpp:6164-6230) comes from ElasTriOrder:
//    from counting sorts for distinct integer points, or, for other sets,
//    after Triangle's own vertexsort() and duplicate removal, which decide
//    which copy of a repeated point survives.  For distinct points the order
//    is unique: both routines compare by (axis, other axis), a strict total
//    order, each cut keeps the `divider` smallest points on its left, and
//    every 2- or 3-point leaf ends sorted by (x, y).  Their random pivots only
//    change how fast the order is reached, and the divide-and-conquer path
//    draws no other random numbers.
//
//  * The merge runs split over threads.  divconqrecurse() on n vertices
//    triangulates the first n/2 of the order, then the rest, then merges the
//    two (src/triangle.cpp:6092-6104), so the triangles are allocated in
//    post-order: left subtree, right subtree, the merge's own; only
//    removeghosts() frees triangles, after the whole merge.  The top `depth`
//    levels of that recursion are cut out: every node of the cut gets its
//    own triangle pool (sharing the mesh's "outer space" triangle, which the
//    merge never writes), the leaves run divconqrecurse() concurrently, an
//    internal node runs mergehulls() as soon as both its children are done
//    (on the thread that finished the second), and the triangles are read
//    out pool by pool in
//    post-order, each pool in allocation order, skipping the dead ghosts --
//    the order in which writeelements() meets them in the single pool of a
//    serial run.  Vertex handles, orientations and the merge arithmetic are
//    Triangle's own, so the triangle list, numbering included, is that of
//    triangulate().  Triangle's namespace-scope state (the exact-arithmetic
//    bounds, the RNG) is written only by triangleinit() on the calling
//    thread, before any task starts, and only read during the merge.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <vector>

#include "elas_triangle.h"

// Triangle allocates only in trimalloc() and frees only in trifree()
// (src/triangle.cpp:1225-1240); both go to the block cache.
#define malloc(bytes) elas_tri_block_alloc(bytes)
#define free(ptr) elas_tri_block_free(ptr)
#include "triangle.cpp"
#undef malloc
#undef free

namespace TRILIB {

namespace {

// triangulate()'s steps up to the divide-and-conquer (src/triangle.cpp:
// 8503-8530 and delaunay(), :6237-6250) for n >= 3 points.  transfernodes()
// (:7638-7700) is taken apart for points without attributes or markers: its
// n poolalloc() calls take the first n items of the vertex pool's first
// block, which holds them all (no item is dead yet), so they are taken here
// at once; vertex i is the i-th item.  Filling them (coordinates, mark,
// type) and the bounding box are left to elas_fill and the caller, so the
// fill can run on the threads that use the vertices.
vertex elas_begin(struct mesh& mesh, struct behavior& b, int n) {
    struct mesh* m = &mesh;  // Triangle's vertex macros name the mesh `m`
    char sw[] = "zQB";
    char* swp = sw;
    triangleinit(m);
    parsecommandline(1, &swp, &b);
    m->steinerleft = b.steiner;
    m->invertices = n;
    m->mesh_dim = 2;
    m->nextras = 0;
    m->readnodefile = 0;
    b.weighted = 0;
    initializevertexpool(m, &b);
    struct memorypool* vp = &m->vertices;
    const vertex base = (vertex)vp->nextitem;
    vp->nextitem = (int*)((char*)vp->nextitem + (size_t)n * (size_t)vp->itembytes);
    vp->unallocateditems -= n;
    vp->maxitems += n;
    vp->items += n;
    m->eextras = 0;
    initializetrisubpools(m, &b);
    return base;
}

// input point i's vertex, as transfernodes() fills it
inline vertex elas_vertex(struct mesh* m, vertex base, int i) {
    return (vertex)((char*)base + (size_t)i * (size_t)m->vertices.itembytes);
}
inline void elas_fill(struct mesh* m, vertex v, const float* p) {
    v[0] = p[0];
    v[1] = p[1];
    setvertexmark(v, 0);
    setvertextype(v, INPUTVERTEX);
}

// transfernodes()'s bounding box and sweepline flag (:7684-7696)
void elas_bounds(struct mesh* m, float xmin, float xmax, float ymin, float ymax) {
    m->xmin = xmin;
    m->xmax = xmax;
    m->ymin = ymin;
    m->ymax = ymax;
    m->xminextreme = 10 * m->xmin - 9 * m->xmax;
}

struct SplitNode {
    int lo, size, axis, level;
    int left, right;  // child nodes, -1 for a leaf
    int parent;       // -1 for the root
    struct mesh sub;  // the node's own triangle pool (the root uses the caller's mesh)
    struct otri farleft, farright;
};

// A run of a node's triangle pool for the readout: items [first, first +
// count) of the pool's first block (when it holds them all), or the whole
// pool (first -1).
struct ReadChunk {
    int node, first, count;
    int live;    // its live triangles
    size_t off;  // their place in the element list (ints)
};
#define ELAS_TRI_READ_CHUNK 2048

struct SplitRun {
    struct mesh* m;
    struct behavior* b;
    vertex* vert;       // vertices by the order's point index, or (null) input point i's vertex
    vertex base;        //   is item i of the pool, filled by the leaf that holds it from
    const float* pts;   //   the input points
    vertex* sortarray;  // the divide-and-conquer order (filled by the leaves)
    ElasTriOrder* order;
    std::vector<SplitNode> nodes;
    int root;
    std::vector<int> todo;  // the leaves
    std::vector<ReadChunk> chunks;  // the readout, in element-list order
    std::unique_ptr<std::atomic<int>[]> arrived;  // per node: children done
    int32_t* dst;
    bool dst16;  // dst takes 16-bit corner indices
    // Vertex numbers.  numbernodes() (src/triangle.cpp:7781-7796) numbers
    // the vertices in pool order, and here that is their allocation order:
    // no vertex is ever freed (vertexdealloc() has no caller) and "zQB" does
    // not jettison duplicates.  initializevertexpool() places every input
    // vertex in the pool's first block, so a vertex's number is its offset
    // from the first one over the item size: with a power-of-two item size
    // (16 bytes here) the read shifts by vshift, else (vshift -1) it reads
    // the marks numbernodes() wrote.
    const char* vbase;
    int vshift;
};

// post-order build of the cut: nodes of more than 3 vertices above `depth`
// split as divconqrecurse() (and ElasTriOrder::cut) split them
int split_build(std::vector<SplitNode>& nodes, int lo, int size, int axis, int level, int depth) {
    SplitNode nd;
    nd.lo = lo;
    nd.size = size;
    nd.axis = axis;
    nd.level = level;
    nd.left = nd.right = nd.parent = -1;
    if (depth > 0 && size > 3) {
        const int divider = size >> 1;
        nd.left = split_build(nodes, lo, divider, 1 - axis, level + 1, depth - 1);
        nd.right = split_build(nodes, lo + divider, size - divider, 1 - axis, level + 1, depth - 1);
    }
    nodes.push_back(std::move(nd));
    const int me = (int)nodes.size() - 1;
    if (nodes[me].left >= 0) nodes[nodes[me].left].parent = nodes[nodes[me].right].parent = me;
    return me;
}

struct mesh* node_mesh(SplitRun* r, int i) { return i == r->root ? r->m : &r->nodes[i].sub; }

// a node's own pool: the mesh's sizes and flags, its dummytri, room for about
// two triangles per vertex in the first block
void node_pool(SplitRun* r, int i) {
    if (i == r->root) return;
    SplitNode& nd = r->nodes[i];
    nd.sub = *r->m;
    poolinit(&nd.sub.triangles, r->m->triangles.itembytes, TRIPERBLOCK, 2 * nd.size + 8,
             r->m->triangles.alignbytes);
}

// an internal node: mergehulls() of its children's triangulations, as the
// tail of divconqrecurse() does it
void split_merge(SplitRun* r, int i) {
    SplitNode& nd = r->nodes[i];
    node_pool(r, i);
    struct otri farleft = r->nodes[nd.left].farleft, innerleft = r->nodes[nd.left].farright;
    struct otri innerright = r->nodes[nd.right].farleft, farright = r->nodes[nd.right].farright;
    mergehulls(node_mesh(r, i), r->b, &farleft, &innerleft, &innerright, &farright, nd.axis);
    nd.farleft = farleft;
    nd.farright = farright;
}

// a leaf: the cuts below the split, its part of the order, its
// triangulation; then up the tree, each merge run by the thread that
// finished the second of its two children (the first one stops there), so a
// merge starts as soon as both halves exist.  The counter's acquire-release
// increment hands the first child's hull and pool over to the merging thread.
void split_leaf(void* arg, int t) {
    SplitRun* r = (SplitRun*)arg;
    const int i = r->todo[t];
    SplitNode& nd = r->nodes[i];
    r->order->cut(nd.lo, nd.size, nd.axis, -1);
    const int32_t* arr = r->order->arr.data();
    if (r->vert) {
        for (int k = nd.lo; k < nd.lo + nd.size; k++) r->sortarray[k] = r->vert[arr[k]];
    } else {
        for (int k = nd.lo; k < nd.lo + nd.size; k++) {
            const int i = arr[k];
            const vertex v = elas_vertex(r->m, r->base, i);
            elas_fill(r->m, v, r->pts + 2 * (size_t)i);
            r->sortarray[k] = v;
        }
    }
    node_pool(r, i);
    divconqrecurse(node_mesh(r, i), r->b, r->sortarray + nd.lo, nd.size, nd.axis, &nd.farleft, &nd.farright);
    for (int j = nd.parent; j >= 0; j = r->nodes[j].parent) {
        if (r->arrived[j].fetch_add(1, std::memory_order_acq_rel) == 0) break;
        split_merge(r, j);
    }
}

// A chunk's triangles in allocation order.  A pool whose items all lie in
// its first block (the usual case) is read in chunks of item indices, each
// visited as traverse() visits it: from the first block's first aligned
// item (traversalinit(), :1786-1795) in steps of the item size.
template <class F>
void chunk_walk(SplitRun* r, const ReadChunk& ch, F&& f) {
    struct memorypool* pool = &node_mesh(r, ch.node)->triangles;
    if (ch.first >= 0) {
        const unsigned long long alignptr = (unsigned long long)(pool->firstblock + 1);
        char* item0 = (char*)(alignptr + (unsigned long long)pool->alignbytes - (alignptr % (unsigned long long)pool->alignbytes));
        for (int j = ch.first; j < ch.first + ch.count; j++) f((triangle*)(item0 + (size_t)j * (size_t)pool->itembytes));
    } else {
        traversalinit(pool);
        triangle* t;
        while ((t = (triangle*)traverse(pool)) != (triangle*)NULL) f(t);
    }
}

// the live triangles of a chunk
void split_count(void* arg, int t) {
    SplitRun* r = (SplitRun*)arg;
    ReadChunk& ch = r->chunks[t];
    int live = 0;
    chunk_walk(r, ch, [&](triangle* tp) { live += !deadtri(tp); });
    ch.live = live;
}

// a chunk's live triangles into the element list at its offset, corners as
// writeelements() numbers them (src/triangle.cpp:7804-7860, order 1)
void split_write(void* arg, int t) {
    SplitRun* r = (SplitRun*)arg;
    struct mesh* m = r->m;  // Triangle's vertex macros name the mesh `m`
    const ReadChunk& ch = r->chunks[t];
    int32_t* d32 = r->dst + ch.off;
    uint16_t* d16 = (uint16_t*)r->dst + ch.off;
    size_t k = 0;
    struct otri tri;
    tri.orient = 0;
    vertex p[3];
    chunk_walk(r, ch, [&](triangle* tp) {
        if (deadtri(tp)) return;
        tri.tri = tp;
        org(tri, p[0]);
        dest(tri, p[1]);
        apex(tri, p[2]);
        for (int c = 0; c < 3; c++, k++) {
            const int32_t id = r->vshift >= 0 ? (int32_t)(((const char*)p[c] - r->vbase) >> r->vshift) : vertexmark(p[c]);
            if (r->dst16)
                d16[k] = (uint16_t)id;
            else
                d32[k] = id;
        }
    });
}

// The merge over `vert` (count vertices; order->lx / ly hold their sort
// orders; with vert null, the unfilled input vertices from base, filled from
// pts by the leaves) cut into at most 2^depth subtrees, the element list
// into *out, or into dst when it holds cap triangles (as 16-bit indices with
// dst16); returns ntri.
int elas_finish_split(struct mesh& mesh, struct behavior& b, vertex* vert, vertex base, const float* pts, int count,
                      ElasTriOrder* order, int depth, ElasParallelFor par, void* parctx, std::vector<int32_t>* out,
                      int32_t* dst, int64_t cap, bool dst16, bool topcut_done) {
    SplitRun r;
    r.m = &mesh;
    r.b = &b;
    r.vert = vert;
    r.base = base;
    r.pts = pts;
    r.order = order;
    r.sortarray = (vertex*)trimalloc(count * (int)sizeof(vertex));
    r.nodes.reserve(2u << std::min(depth, 12));
    r.root = split_build(r.nodes, 0, count, 0, 0, depth);
    r.dst = nullptr;
    r.dst16 = false;
    // the top of the order (the cuts the split nodes make; already made when
    // the caller built the order), then the leaves, each followed by the
    // merges its completion unlocks, up to the root
    if (!topcut_done) order->cut(0, count, 0, depth);
    r.arrived.reset(new std::atomic<int>[r.nodes.size()]);
    for (size_t i = 0; i < r.nodes.size(); i++) r.arrived[i].store(0, std::memory_order_relaxed);
    r.todo.clear();
    for (int i = 0; i < (int)r.nodes.size(); i++)
        if (r.nodes[i].left < 0) r.todo.push_back(i);
    par(parctx, (int)r.todo.size(), split_leaf, &r);
    trifree((int*)r.sortarray);
    removeghosts(&mesh, &b, &r.nodes[r.root].farleft);
    traversalinit(&mesh.vertices);
    r.vbase = (const char*)mesh.vertices.pathitem;
    r.vshift = -1;
    const int ib = mesh.vertices.itembytes;
    if (ib > 0 && (ib & (ib - 1)) == 0 && mesh.vertices.maxitems <= mesh.vertices.itemsfirstblock)
        for (r.vshift = 0; (1 << r.vshift) < ib;) r.vshift++;
    else
        numbernodes(&mesh, &b);
    // the pools in post-order (the nodes' order), each in allocation order
    r.chunks.clear();
    for (int i = 0; i < (int)r.nodes.size(); i++) {
        const struct memorypool* pool = &node_mesh(&r, i)->triangles;
        if (pool->maxitems <= pool->itemsfirstblock) {
            const int items = (int)pool->maxitems;
            for (int f = 0; f < items; f += ELAS_TRI_READ_CHUNK)
                r.chunks.push_back(ReadChunk{i, f, std::min(ELAS_TRI_READ_CHUNK, items - f), 0, 0});
        } else {
            r.chunks.push_back(ReadChunk{i, -1, 0, 0, 0});
        }
    }
    // count, place, then write each chunk's triangles where they belong
    par(parctx, (int)r.chunks.size(), split_count, &r);
    size_t total = 0;
    for (ReadChunk& ch : r.chunks) {
        ch.off = total;
        total += 3 * (size_t)ch.live;
    }
    const int nt = (int)(total / 3);
    if (dst) {
        if (nt <= cap) {
            r.dst = dst;
            r.dst16 = dst16;
            par(parctx, (int)r.chunks.size(), split_write, &r);
        }
    } else {
        out->resize(total);
        r.dst = out->data();
        par(parctx, (int)r.chunks.size(), split_write, &r);
    }
    for (int i = 0; i < (int)r.nodes.size(); i++)
        if (i != r.root) pooldeinit(&r.nodes[i].sub.triangles);
    triangledeinit(&mesh, &b);
    return nt;
}

}  // namespace

namespace {

struct Setup {
    struct mesh mesh;
    struct behavior b;
    vertex base;  // input point i's vertex is item i of the vertex pool
    const float* pts;
    int n;
};

}  // namespace

// Triangle's setup of a "zQB" triangulation of n >= 3 points [x0, y0, ...]
// (any thread; the vertices are taken, not yet filled); elas_tri_finish()
// takes it over.
void* elas_tri_setup(const float* pts, int n) {
    Setup* st = new Setup();
    st->base = elas_begin(st->mesh, st->b, n);
    st->pts = pts;
    st->n = n;
    return st;
}

// The triangulation, corners numbered by input position, into *out (or dst,
// see elas_finish_split).  `order` is scratch; with `integer` set it already
// holds the input points' sort orders and bounds (elas_tri_order_integer)
// and its top `depth` levels of cuts, and the leaves fill the vertices, else
// they are filled here, then Triangle's vertexsort() and duplicate removal
// (:6176-6198) run and the order is built over the survivors.  Returns ntri,
// or -1 when a single distinct point is left: Triangle's divconqrecurse() has
// no case for one vertex and recurses until the stack overflows (two
// distinct points give an empty mesh, as in Triangle).
int elas_tri_finish(void* setup, ElasTriOrder* order, bool integer, int depth, ElasParallelFor par, void* parctx,
                    std::vector<int32_t>* out, int32_t* dst, int64_t cap, bool dst16) {
    Setup* st = (Setup*)setup;
    struct mesh* m = &st->mesh;  // Triangle's vertex macros name the mesh `m`
    const int n = st->n;
    const float* pts = st->pts;
    int count = n, nt = -1;
    vertex* byinput = nullptr;
    if (integer) {
        // integer coordinates below 2^24: the bounds are exact as floats
        elas_bounds(m, (float)order->lohi[0], (float)order->lohi[1], (float)order->lohi[2], (float)order->lohi[3]);
    } else {
        byinput = (vertex*)trimalloc(n * (int)sizeof(vertex));
        float xmin = pts[0], xmax = pts[0], ymin = pts[1], ymax = pts[1];
        for (int i = 0; i < n; i++) {
            const vertex v = elas_vertex(m, st->base, i);
            elas_fill(m, v, pts + 2 * (size_t)i);
            byinput[i] = v;
            const float x = pts[2 * i], y = pts[2 * i + 1];
            xmin = (x < xmin) ? x : xmin;
            xmax = (x > xmax) ? x : xmax;
            ymin = (y < ymin) ? y : ymin;
            ymax = (y > ymax) ? y : ymax;
        }
        elas_bounds(m, xmin, xmax, ymin, ymax);
        vertexsort(byinput, n);
        int i = 0;
        for (int j = 1; j < n; j++) {
            if (byinput[i][0] == byinput[j][0] && byinput[i][1] == byinput[j][1]) {
                setvertextype(byinput[j], UNDEADVERTEX);
                m->undeads++;
            } else {
                i++;
                byinput[i] = byinput[j];
            }
        }
        count = i + 1;
        if (count >= 2) {
            std::vector<float> xy(2 * (size_t)count);
            for (int k = 0; k < count; k++) {
                xy[2 * k] = byinput[k][0];
                xy[2 * k + 1] = byinput[k][1];
            }
            elas_tri_order_sorted(xy.data(), count, *order);
        }
    }
    if (count >= 2) {
        nt = elas_finish_split(st->mesh, st->b, byinput, st->base, pts, count, order, depth, par, parctx, out, dst,
                               cap, dst16, integer);
    } else {
        triangledeinit(&st->mesh, &st->b);
        if (out) out->clear();
    }
    if (byinput) trifree((int*)byinput);
    delete st;
    return nt;
}

}  // namespace TRILIB

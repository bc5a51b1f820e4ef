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
int elas_triangulate_dedup(const float* pts, int n, bool (*arrange)(const float* xy, int count, int* arr),
                           int** trilist) {
    struct mesh mesh;
    struct mesh* m = &mesh;  // Triangle's vertex macros name the mesh `m`
    struct behavior b;
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
        *trilist = (int*)NULL;
        return -1;
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
    return elas_finish(mesh, b, sortarray, count, trilist);
}

}  // namespace TRILIB

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
// passes that order in.  For distinct points the order is unique: both
// routines compare by (axis, other axis), a strict total order, each cut
// keeps the `divider` smallest points on its left, and every 2- or 3-point
// leaf ends sorted by (x, y).  Their random pivots only change how fast the
// order is reached, and the divide-and-conquer path draws no other random
// numbers, so the mesh and its triangle numbering are those of triangulate().
#include "triangle.cpp"

namespace TRILIB {

// pts: [n][2], n >= 3, distinct; arr: the divide-and-conquer vertex order as
// input indices.  *trilist receives a malloc'ed [ntri][3] zero-based list
// (free() it); returns ntri.
int elas_triangulate_arranged(const float* pts, int n, const int* arr, int** trilist) {
    struct mesh m;
    struct behavior b;
    char sw[] = "zQB";
    char* swp = sw;
    triangleinit(&m);
    parsecommandline(1, &swp, &b);
    m.steinerleft = b.steiner;
    transfernodes(&m, &b, (float*)pts, (float*)NULL, (int*)NULL, n, 0);
    m.eextras = 0;
    initializetrisubpools(&m, &b);
    vertex* byinput = (vertex*)trimalloc(n * (int)sizeof(vertex));
    vertex* sortarray = (vertex*)trimalloc(n * (int)sizeof(vertex));
    traversalinit(&m.vertices);
    for (int i = 0; i < n; i++) byinput[i] = vertextraverse(&m);
    for (int k = 0; k < n; k++) sortarray[k] = byinput[arr[k]];
    trifree((int*)byinput);
    struct otri hullleft, hullright;
    divconqrecurse(&m, &b, sortarray, n, 0, &hullleft, &hullright);
    trifree((int*)sortarray);
    const long hull = removeghosts(&m, &b, &hullleft);
    m.hullsize = m.triangles.items == 0 ? 0l : hull;
    // triangulate() numbers the vertices in writenodes(); without the point
    // list numbernodes() gives the same numbers (no -j: every vertex counts).
    numbernodes(&m, &b);
    float* attr = (float*)NULL;
    *trilist = (int*)NULL;
    writeelements(&m, &b, trilist, &attr);
    const int nt = (int)m.triangles.items;
    triangledeinit(&m, &b);
    return nt;
}

}  // namespace TRILIB

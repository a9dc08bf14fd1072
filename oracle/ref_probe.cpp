// Builds the reference engine (cpp/nbody.cpp, compiled from its own source
// where it lies) together with one accessor that copies out the tree the last
// nbody_accel_active call left in the reference's global state.  TEST
// INFRASTRUCTURE ONLY: the recipe in oracle/Makefile writes the result to
// oracle/_ref/, and only tests/golden/make_golden.py and tests load it.
#include "nbody.cpp"

extern "C" int ref_tree_export(int* start, int* count, int* next, int* leaf, float* com4,
                               float* open_r, int* idx, int* groups, int* ngroups) {
    const Tree& T = g_tree;
    const int nn = int(T.nodes.size());
    for (int k = 0; k < nn; ++k) {
        const Node& q = T.nodes[k];
        if (start) start[k] = q.start;
        if (count) count[k] = q.count;
        if (next) next[k] = q.next;
        if (leaf) leaf[k] = q.leaf;
        if (com4) { com4[4 * k] = q.cx; com4[4 * k + 1] = q.cy; com4[4 * k + 2] = q.cz; com4[4 * k + 3] = q.m; }
        if (open_r) open_r[k] = q.open_r;
    }
    if (idx) for (size_t i = 0; i < T.idx.size(); ++i) idx[i] = T.idx[i];
    if (groups) for (size_t g = 0; g < T.groups.size(); ++g) groups[g] = T.groups[g];
    if (ngroups) *ngroups = int(T.groups.size());
    return nn;
}

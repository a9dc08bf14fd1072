// SPH on the device (gf_sph.hip): exact k-nearest-neighbour solve, smoothing
// lengths, density and pressure/viscosity forces of the gas.
#pragma once
#include "gf_tree.h"

struct GfSph {
    GfTree tree;                 // octree of the gas particles
    int cap_n = 0, cap_k = 0;
    // per gas particle, gas-tree order
    double *vx = nullptr, *vy = nullptr, *vz = nullptr, *u = nullptr;
    double *r = nullptr;         // search radius of the current pass
    double *r_lo = nullptr, *r_hi = nullptr;   // radius bracket of the iteration
    double *dk = nullptr;        // distance to the k-th neighbour (self included)
    double *h = nullptr, *rho = nullptr;
    double *rec = nullptr;       // 16 doubles per particle: packed neighbour records
    double *p4 = nullptr;        // (x, y, z, -) per particle for the neighbour walks
    double *knode = nullptr;     // 8 doubles per node: packed nodes for the neighbour walks
    int *gmark = nullptr;        // per group: pass tag of its last retry request
    int *retry = nullptr;        // two lists of groups that need another pass
    int8_t *done = nullptr;
    int *passes = nullptr;
    int *inv = nullptr;          // input index -> tree position
    int *leaf_of = nullptr;      // leaf node of each tree position
    int *anc = nullptr;          // per group: its node and ancestors, GF_ANC ints (-1 past the root)
    int *items = nullptr;        // work items of a pass from a caller's radii: int4 (first group
                                 // position, groups, first particle, members) per item
    int *nb = nullptr;           // [k][n] neighbours: tree positions (or input indices)
    int *state = nullptr;        // pass counters (layout in gf_sph.hip)
    int *state_host = nullptr;   // their copy at the end of a solve (pinned host memory)
    int *ovf = nullptr, *ovf2 = nullptr;
    int *gstack = nullptr;
    int cap_groups = 0;
    int fast_blocks = 0;         // persistent blocks of the fast neighbour pass (0 = default)
    int merge_members = 0;       // members of a merged work item at most (0 = default, 1 = none)
    int pair_f32 = 0;            // density and forces: FP32 pair arithmetic, FP64 sums
    int last_passes = 0;         // pass launches of the last solve
    long long stat_queued = 0;   // groups the fast pass handed on, summed over passes
    long long stat_queued2 = 0;  // groups handed on to the global-stack pass
    // the solve in flight (gf_sph_knn_begin .. gf_sph_knn_end)
    int knn_kq = 0, knn_pos_out = 0, knn_sorted = 0, knn_narrow = 0, knn_items = 0;
    int knn_enqueued = 0;        // pass launches enqueued so far
    int knn_open = 0;            // begun, not yet ended
};

void gf_sph_free(GfSph& S);

// Persistent blocks the fast neighbour pass launches from a caller's first
// radius (S.fast_blocks, or the default of one per resident slot of the
// narrow shape: 1280).  From the cell densities the default is 1024.
int gf_sph_fast_blocks(const GfSph& S);
// Members a work item of the first pass from a caller's radii holds at most
// (S.merge_members, or the default); 1 = every group is an item of its own.
int gf_sph_merge_members(const GfSph& S);

// Builds the gas tree and gathers velocities and energies into tree order.
// pos, vel (n, 3), mass, u (n) are device arrays in input order (vel and u
// may be null when only neighbours are wanted).
int gf_sph_setup(GfSph& S, int n, const double* pos, const double* vel, const double* mass,
                 const double* u, hipStream_t s);
// The same for the first n particles of the arrays the tree T was last
// built from: the gas tree is T's sorted order restricted to them
// (gf_tree_build_subset), without a sort of its own.
int gf_sph_setup_from_tree(GfSph& S, const GfTree& T, int n, const double* pos, const double* vel,
                           const double* mass, const double* u, hipStream_t s);

// Exact kq nearest neighbours of every gas particle (the particle itself
// included) by radius iteration: a pass gathers, per particle, every gas
// particle within its search radius into a max-heap of (d^2, input index);
// particles with fewer than kq inside grow their radius and take another
// pass.  r_init (input order, may be null) is the first radius; without it
// the radius comes from the density of the smallest tree cell holding >= kq
// particles.  tree_pos_out: neighbours as tree positions (else input
// indices); sorted: each row ascending in (d^2, index).
int gf_sph_knn(GfSph& S, int kq, const double* r_init, bool tree_pos_out, bool sorted,
               int max_passes, hipStream_t s);
// The solve in two halves: _begin enqueues the set-up and `passes` pass
// launches and returns without waiting (later passes find nothing to do once
// every particle is resolved); _end waits for them, checks the device's
// error flags and enqueues further passes, two at a time, until every
// particle is resolved or max_passes launches were made.  Nothing may use
// the neighbour table between the two calls.
int gf_sph_knn_begin(GfSph& S, int kq, const double* r_init, bool tree_pos_out, bool sorted,
                     int passes, hipStream_t s);
int gf_sph_knn_end(GfSph& S, int max_passes, hipStream_t s);
// Enqueues `passes` more passes of the solve in flight (after a _begin with
// none, the set-up alone, a caller can order other streams after it).
int gf_sph_knn_more(GfSph& S, int passes, hipStream_t s);

// h = max(d_k / 2, eps / 2) (galaxy_formation/simulation.py:84) from the
// k-th neighbour distance of the last solve.
int gf_sph_smoothing(GfSph& S, int k, double eps, hipStream_t s);

// Density and forces (galaxy_formation/sph.py:57-106) over the tree-order
// neighbour table of the last solve (k columns).  Outputs are written in
// input order: rho, dudt, dt, divv (n) and acc (n, 3); any may be null.
int gf_sph_forces_tree(GfSph& S, int k, double gamma, double alpha, double beta, double* rho,
                       double* acc, double* dudt, double* dt, double* divv, double* h_out,
                       hipStream_t s);

// Density and forces over a caller's neighbour table nb (n, k) int64 of
// input indices, all arrays device, input order (sph.py:57-106 exactly).
int gf_sph_forces_table(int n, int k, const double* pos, const double* vel, const double* mass,
                        const int64_t* nb, const double* h, const double* u, double gamma,
                        double alpha, double beta, double* rho, double* acc, double* dudt,
                        double* dt, double* divv, int* scratch_nb, hipStream_t s);

// Bytes of device scratch gf_sph_forces_table needs.
size_t gf_sph_table_scratch_bytes(int n, int k);

// Writes the neighbour table of the last solve as (n, k) int64 input indices
// and (n, k) distances, rows in input order (requires a sorted solve).
int gf_sph_export_knn(GfSph& S, int kq, int64_t* nb, double* dist, hipStream_t s);

// Number density of the smallest gas-tree cell holding >= kq particles
// around each gas particle, and (out_count, may be null) that cell's particle
// count (input order, device arrays of n).
int gf_sph_cell_density(GfSph& S, int kq, double* out, double* out_count, hipStream_t s);
// d_k of the last solve, input order (device array of n).
int gf_sph_dk_out(GfSph& S, double* dk, hipStream_t s);
// Tree-position table of the last solve as (n, k) int64 input indices, rows
// in input order, neighbours unordered (device array).
int gf_sph_table_out(GfSph& S, int k, int64_t* nb, hipStream_t s);

// knn_from_candidates (sph.py:109-135): per row, the k nearest of the m
// candidates (-1 = none), ascending (d^2, index).  *ok = 1 when every row had
// k valid candidates.  Device arrays.
int gf_knn_from_candidates(int n, int m, const double* pos, const int64_t* cand, int k,
                           int64_t* nb, double* dist, int* ok, hipStream_t s);

// W(r, h) and (1/r) dW/dr (sph.py:31-54) for n (r, h) pairs.
int gf_sph_kernel_eval(int n, const double* r, const double* h, double* w, double* g,
                       hipStream_t s);

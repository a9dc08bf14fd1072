// Device-resident neighbour cache of sph.NeighbourFinder (gf_verlet.hip).
#pragma once
#include <algorithm>
#include <vector>

#include "gf_common.h"

// Candidates per cache row: one per lane of the re-selection's wave.
#define GF_VERLET_MAX_CAND 64

struct GfVerlet {
    int n_total = 0;             // particles of the simulation at the last full search
    int rows = 0;                // gas particles then (cache rows)
    int m = 0;                   // candidates per row
    int* owner = nullptr;        // n_total: cache row of each particle, -1 = was not gas
    int* cand = nullptr;         // rows x m: candidates as global particle indices
    int* loc = nullptr;          // n_total: current local index, -1 between calls
    int* flags = nullptr;        // [0] ok, [1] bad gas index, [2] bad neighbour index
    size_t cap_owner = 0, cap_cand = 0, cap_loc = 0, cap_flags = 0;
};

// Replaces the cache with the rows of a full search: g (n, device int64) the
// global indices of the gas, nb (n x m, device int64) each row's m nearest as
// local indices.  Waits.
int gf_verlet_store(GfVerlet& V, int n_total, int n, int m, const int64_t* g, const int64_t* nb,
                    hipStream_t s);
// Re-selects the k nearest of each current gas particle (g, pos: device) among
// its cached candidates that are still gas, ascending in (d^2, local index),
// into nb (n x k int64 local) and dist.  Returns 1 when every row was served,
// 0 when the cache cannot serve this gas list (then the outputs are
// incomplete), -1 on an error.  row_scratch: n ints of device scratch.  Waits.
int gf_verlet_reselect(GfVerlet& V, int n_total, int n, int k, const int64_t* g,
                       const double* pos, int* row_scratch, int64_t* nb, double* dist,
                       hipStream_t s);
// The cache as host arrays: owner (n_total), cand (rows x m) global indices
// (either may be null).  Returns the row count.
int gf_verlet_export(const GfVerlet& V, int64_t* owner, int64_t* cand, hipStream_t s);
// nb, dist (n x k) = the first k columns of src_nb, src_dist (n x m), device.
int gf_first_columns(int n, int m, int k, const int64_t* src_nb, const double* src_dist,
                     int64_t* nb, double* dist, hipStream_t s);
void gf_verlet_free(GfVerlet& V);

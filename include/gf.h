/*
 * libgf.so: C ABI of the MI355X (gfx950) force path.
 *
 * Conventions: every function returns 0 (or a documented non-negative value)
 * on success and a negative value on failure; gf_last_error() then holds the
 * message.  Arrays are C-contiguous: positions, velocities and accelerations
 * are (n, 3) row-major float64; masses, smoothing lengths, energies (n,)
 * float64; neighbour tables (n, k) int64 of local indices.  Host-pointer
 * entry points copy their inputs to the GPU, run on the library's stream and
 * return when the outputs are back in host memory.  *_dev entry points take
 * device pointers and a hipStream_t (NULL = HIP's default stream) and return
 * without synchronising.  The library keeps state between calls (the last
 * gravity tree, staging buffers), like the reference's global tree
 * (cpp/nbody.cpp:307), and is not re-entrant.
 *
 * File:line citations are in the reference HugoGW/Galaxy_formation.
 */
#ifndef GF_H
#define GF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------ the reference's engine ABI
 * Same names and arguments as cpp/nbody.cpp:415-505, bound with the argtypes
 * of galaxy_formation/engine.py:26-35.  The reference returns void; these
 * return int status, which ctypes' default restype (c_int) reads unchanged. */

/* Replaces nbody_num_threads (cpp/nbody.cpp:417): the GPU's compute units. */
int nbody_num_threads(void);

/* Replaces nbody_accel_active (cpp/nbody.cpp:422-425): Barnes-Hut
 * accelerations (G = 1, Plummer eps) of the active particles (active[i] != 0;
 * NULL = all), written to acc; other rows are left untouched.  rebuild = 0
 * refits the tree of the previous call (a change of n forces a build).  The
 * interaction precision is the one gf_set_precision selected. */
int nbody_accel_active(int n, const double* pos, const double* mass, double* acc, double theta,
                       double eps, const uint8_t* active, int rebuild);

/* Replaces nbody_accel_direct (cpp/nbody.cpp:428-441): FP64 O(n^2) sum. */
int nbody_accel_direct(int n, const double* pos, const double* mass, double* acc, double eps);

/* Replaces nbody_render (cpp/nbody.cpp:445-469): orthographic projection and
 * bilinear deposit into img (ngroups, H, W) float32. */
int nbody_render(int n, const double* pos, const uint8_t* group, int ngroups, const double* rot,
                 const double* center, double scale, int W, int H, float* img);

/* Replaces nbody_tonemap (cpp/nbody.cpp:474-503): RGB8 image (H, W, 3). */
int nbody_tonemap(int W, int H, int ngroups, const float* img, const float* glow, int gw, int gh,
                  const float* colors, float inv_ref, float K, float exposure, uint8_t* out);

/* --------------------------------------------------------------- library */

/* Message of the last failure. */
const char* gf_last_error(void);
/* ABI version of this header. */
int gf_version(void);
/* Device name, architecture and compute-unit count as text. */
int gf_device_info(char* buf, int len);
/* Interaction precision of nbody_accel_active: 0 = FP32 pair arithmetic on
 * group-relative coordinates with compensated accumulation (default), 1 =
 * FP64, 2 = as 0 with each accepted node's traceless quadrupole moment added
 * to its monopole (no counterpart in the reference, which evaluates
 * monopoles: cpp/nbody.cpp:386-408).  Tree, centres of mass, quadrupoles and
 * acceptance tests are FP64 in all three; the interaction set is the same. */
int gf_set_precision(int precision);
int gf_get_precision(void);
/* Precision 2 only, no counterpart in the reference: an accepted node whose
 * quadrupole term is estimated (3 max |Q_ij| / (m d^2), d the distance from
 * its centre of mass to the nearest point of the group's sphere) below tol
 * of its monopole is evaluated as a monopole.  Default 0.01; 0 evaluates
 * every accepted node with its quadrupole.  tol in [0, 1]. */
int gf_set_quad_tolerance(double tol);
double gf_get_quad_tolerance(void);
/* Precision 2 only, no counterpart in the reference, 0 = off (the default):
 * the acceleration rows a gravity call receives in acc (the caller's
 * previous step, as simulation.py:64 passes self.acc_grav) set the rule
 * instead.  An accepted node whose quadrupole term is bounded (2.858 |Q|_F /
 * d^4, d as above) below tol times the smallest |acc| of the group's active
 * members is evaluated as a monopole.  Groups with an active member whose
 * row is zero or not finite, and groups of more than 64 particles, keep the
 * gf_set_quad_tolerance rule.  tol in [0, 1]. */
int gf_set_quad_relative_tolerance(double tol);
double gf_get_quad_relative_tolerance(void);
/* Pair arithmetic of the device-resident density and forces
 * (gf_hydro_forces_dev; sph.py:57-106 is FP64 under Numba fastmath): 0 = FP64
 * (the default), 1 = FP32 pair terms from FP64 separations, FP64 sums.  The
 * neighbours and h do not depend on it; the density differs by about 1e-7
 * relative.  The host-array sph_forces path (gf_sph_forces) stays FP64. */
int gf_set_sph_precision(int precision);
int gf_get_sph_precision(void);
/* Scheduling, no counterpart in the reference (sph.py:165-166 calls a k-d
 * tree): the number of persistent blocks of the fast neighbour pass.
 * 0 = the default: 1280 from a caller's first radius, every resident slot
 * of that pass (5 blocks of 4 waves per CU), fastest when the solve has the
 * GPU to itself; 1024 from the cell densities (3 resident per CU).  A solve
 * run on a second stream beside the gravity walk leaves the walk more of
 * each CU with fewer (512: the measured step of bench.py).  The neighbours
 * do not depend on it.  blocks in [0, 65536]. */
int gf_set_knn_blocks(int blocks);
int gf_get_knn_blocks(void);
/* Scheduling, no counterpart in the reference: the first pass of a solve
 * from a caller's radii serves runs of sibling gas groups as one work item
 * (one tree walk and one candidate gather) while the item holds at most
 * `members` particles.  0 = the default (40), 1 = every group alone; at
 * most 64.  The neighbours do not depend on it. */
int gf_set_knn_merge(int members);
int gf_get_knn_merge(void);
/* Scheduling, no counterpart in the reference (nbody.cpp:323-411 walks with
 * OpenMP threads): the number of persistent one-wave blocks of the gravity
 * walk, which draws groups from a queue.  0 = the default, 4096 (one per
 * resident slot of the walk on MI355X).  A smaller grid leaves room on each
 * CU for work on another stream.  blocks in [0, 65536]. */
int gf_set_walk_blocks(int blocks);
int gf_get_walk_blocks(void);
/* Testing hook, no counterpart in the reference: lowers the capacity of
 * the gravity walk's LDS stack (walk_stack entries, 1..1024), of the
 * neighbour solve's candidate lists (knn_list, 1..1024) and of the tree
 * build's LDS list of nodes above its bottom-up subtrees (moment_top,
 * 1..2048) so that tests can drive the global-memory, heap and list
 * fallbacks on small inputs; 0 restores the default.  Results do not depend
 * on these capacities. */
int gf_set_test_limits(int walk_stack, int knn_list, int moment_top);
/* Interactions (sum over evaluated particles of their list length) of the
 * last nbody_accel_active call (the count nbody.cpp:397 loops over). */
long long gf_last_interactions(void);

/* FP64 direct sum (cpp/nbody.cpp:428-441) for the nt targets tgt (indices
 * into pos; NULL = all n with nt == n); acc is (nt, 3). */
int gf_accel_direct_subset(int n, const double* pos, const double* mass, int nt,
                           const int64_t* tgt, double* acc, double eps);

/* Node, group and key-depth-cut cell counts of the held tree. */
int gf_tree_info(int* nodes, int* groups, int* deep_cells);
/* Diagnostics, no counterpart in the reference: the number of look-back
 * waits of the single-pass scan (tree builds, gas compactions) that gave up
 * since the library was loaded; 0 unless a scan went wrong. */
int gf_lookback_timeouts(void);
/* The held tree in the layout of cpp/nbody.cpp:38-45 (DFS order): returns the
 * node count (cap = room in the node arrays); any pointer may be NULL.
 * com4 = (cx, cy, cz, m) FP64, box4 = (gx, gy, gz, half) FP64, idx = input
 * index of each tree position (n), groups = group node indices. */
int gf_tree_export(int cap, int* start, int* count, int* next, int* leaf, int* depth,
                   double* com4, double* open_r, double* box4, int* idx, int* groups);

/* Device-pointer gravity (same semantics as nbody_accel_active), with the
 * precision per call. */
int gf_gravity_dev(int n, const double* pos, const double* mass, double* acc, double theta,
                   double eps, const uint8_t* active, int rebuild, int precision, void* stream);
/* The two halves of gf_gravity_dev, so callers can time them apart: build
 * the tree of (pos, mass) (cpp/nbody.cpp:202-256), then walk it for the
 * active particles (cpp/nbody.cpp:335-411). */
int gf_tree_build_dev(int n, const double* pos, const double* mass, double theta, void* stream);
int gf_walk_dev(int n, const uint8_t* active, double* acc, double eps, int precision,
                void* stream);
/* Scheduling, no counterpart in the reference: packs the held tree's walk
 * records for a walk at `precision` (node records level by level, packed
 * particles, and for precision 2 the quadrupoles), which gf_walk_dev
 * otherwise does first.  Called on the walk's stream right after the
 * build, it moves that work beside other work; the walk then reuses the
 * records of the same tree build. */
int gf_walk_prepare_dev(int precision, void* stream);
/* Device-pointer FP64 direct sum for nt targets. */
int gf_direct_dev(int n, const double* pos, const double* mass, int nt, const int64_t* tgt,
                  double* acc, double eps, void* stream);
/* No counterpart in the reference (which reports no energies): the softened
 * FP64 potential phi_i = -sum_{j != i} m_j / sqrt(r_ij^2 + eps^2) of all n
 * particles by direct summation, for the energy of a run (device arrays). */
int gf_potential_dev(int n, const double* pos, const double* mass, double* phi, double eps,
                     void* stream);
/* --------------------------------------------- the caller's step, on the device
 * No counterparts in the reference's C++: these are the NumPy lines of
 * galaxy_formation/simulation.py around the force calls, for a caller that
 * keeps its state on the device (device arrays, FP64, the operations in the
 * order those lines write them). */

/* simulation.py:68-71: acc = acc_grav + acc_sph on the gas rows 0 .. ng-1
 * (acc_grav elsewhere); dt_ind = 0.3 sqrt(eps / max(|acc|, 1e-30)), on the
 * gas rows the smaller of that and dt_c; *dt_grav (one device double) =
 * 0.3 sqrt(eps / max(max |acc|, 1e-30)).  acc_sph and dt_c may be NULL when
 * ng = 0. */
int gf_combine_dev(int n, int ng, double eps, const double* acc_grav, const double* acc_sph,
                   const double* dt_c, double* acc, double* dt_ind, double* dt_grav,
                   void* stream);
/* Sub-step k of a block leapfrog with `rungs` rungs (the drift of every
 * particle, simulation.py:122, 138): pos += dt vel; active[i] = 1
 * where k is a multiple of 2^(rungs-1-rung[i]) (the particles at a boundary
 * of their step), else 0; *count (device int) += the number active. */
int gf_step_drift_dev(int n, double dt, long long k, int rungs, double* pos, const double* vel,
                      const int64_t* rung, uint8_t* active, int* count, void* stream);
/* Half kicks of the active particles (simulation.py:121, 124, 131, 137, 142, 148)
 * by h = dt_max 2^-rung / 2: vel += h acc, and on the gas rows
 * u = max(u + h dudt, u_floor).  Before the kick, mode 1 moves each active
 * particle to the rung of its dt_ind, to a longer step only where sub-step k
 * is also a boundary of it, and mode 2 moves every particle to the rung of its
 * dt_ind (a new ladder); mode 0 keeps the rungs.  The rung of dt is
 * min(max(ceil(log2(max(dt_max / dt, 1))), 0), rungs - 1).  dudt and u may be
 * NULL when ng = 0, dt_ind in mode 0. */
int gf_step_kick_dev(int n, int ng, long long k, int rungs, double dt_max, int mode,
                     const uint8_t* active, const double* acc, const double* dudt,
                     const double* dt_ind, int64_t* rung, double* vel, double* u, double u_floor,
                     void* stream);

/* Interaction count of the last gravity call (synchronises). */
int gf_interactions_dev(long long* out);
/* The part of that count that are node (not particle) entries. */
int gf_node_interactions_dev(long long* out);
/* Waits for the stream and reports kernel-side errors. */
int gf_sync(void* stream);

/* ------------------------------------------------------------------ SPH
 * The reference's SPH is Python + Numba (galaxy_formation/sph.py), with no
 * native ABI; these entry points are what its functions call here. */

/* kernel / kernel_grad (sph.py:31-54) for n (r, h) pairs: w = W(r, h),
 * g = (1/r) dW/dr; either output may be NULL. */
int gf_sph_kernel(int n, const double* r, const double* h, double* w, double* g);

/* sph_forces (sph.py:57-106): density, pressure + viscosity acceleration,
 * du/dt and Courant time step over the neighbour table nb (n, k) int64 of
 * local indices (the particle itself included).  divv (may be NULL) receives
 * the SPH velocity divergence -(1/rho_i) sum_j m_j (v_ij . r_ij) (1/r) dW/dr. */
int gf_sph_forces(int n, int k, const double* pos, const double* vel, const double* mass,
                  const int64_t* nb, const double* h, const double* u, double gamma, double alpha,
                  double beta, double* rho, double* acc, double* dudt, double* dt_i, double* divv);

/* Exact k nearest neighbours of every particle, the particle itself first
 * (replaces the cKDTree query of sph.py:165-166): nb (n, k) int64, dist
 * (n, k), rows ascending in (distance, index).  r_init (may be NULL): first
 * search radius per particle.  Returns the number of passes of the radius
 * iteration; *mean_passes (may be NULL) gets the mean per particle. */
int gf_knn(int n, const double* pos, int k, const double* r_init, int64_t* nb, double* dist,
           double* mean_passes);

/* knn_from_candidates (sph.py:109-135): the k nearest of each row's m
 * candidates (-1 = none), ascending.  Returns 1 when every row had >= k valid
 * candidates, 0 otherwise (rows without enough are left untouched). */
int gf_knn_from_candidates_host(int n, int m, const double* pos, const int64_t* cand, int k,
                                int64_t* nb, double* dist);

/* sph.NeighbourFinder.find (sph.py:151-170), full search: the exact m
 * nearest gas neighbours (the particle itself first) of the n gas particles
 * g (global indices into the simulation's n_total particles) at pos (n, 3),
 * kept on the GPU as candidate lists of global indices; nb (n, k) int64 local
 * indices and dist (n, k) receive the first k, rows ascending in (distance,
 * index).  1 <= k <= m <= min(n, 64).  Replaces the cKDTree query and the
 * cache of sph.py:165-169. */
int gf_nbcache_search(int n_total, int n, const int64_t* g, const double* pos, int m, int k,
                      int64_t* nb, double* dist);
/* sph.NeighbourFinder.find with fresh = False (sph.py:155-163): the k nearest
 * of each gas particle among the candidates of the last search that are
 * still gas (ascending in distance, then local index).  Returns 1 when every
 * particle was served, 0 when the cache cannot serve this gas list (a
 * particle without a cached row, fewer than k candidates left, another
 * n_total): the caller then searches again; outputs are incomplete then. */
int gf_nbcache_reselect(int n_total, int n, const int64_t* g, const double* pos, int k,
                        int64_t* nb, double* dist);
/* Sizes of the held cache (any pointer may be NULL); returns its row count. */
int gf_nbcache_info(int* n_total, int* rows, int* m);
/* The held cache as host arrays: owner (n_total) cache row of each particle
 * (-1 = not gas at the search), cand (rows, m) global indices.  Returns the
 * row count. */
int gf_nbcache_export(int64_t* owner, int64_t* cand);

/* Device-resident hydro of the gas (simulation.py:74-95): gas octree, exact
 * kNN by radius iteration (r_init: first radius per particle, NULL = from
 * the tree's cell densities), h = max(d_k / 2, eps / 2) (simulation.py:84),
 * density and forces (sph.py:57-106).  Inputs and outputs are device arrays
 * in input order; outputs may be NULL.  Returns the number of passes. */
int gf_hydro_dev(int n, const double* pos, const double* vel, const double* mass, const double* u,
                 double eps, int k, const double* r_init, double gamma, double alpha, double beta,
                 double* h, double* rho, double* acc, double* dudt, double* dt, double* divv,
                 void* stream);
/* The three stages of gf_hydro_dev, so callers can time them apart. */
int gf_hydro_setup_dev(int n, const double* pos, const double* vel, const double* mass,
                       const double* u, void* stream);
/* No counterpart in the reference (sph.py:166 builds its own cKDTree): the
 * same set-up for a gas that is the first n rows of pos, the array of the
 * last gf_tree_build_dev, whose Morton order restricted to the gas becomes
 * the gas tree (no bounding box, keys or sort of its own; the gas tree's
 * cells are aligned with the gravity tree's).  The stream waits for that
 * build's sort, which may run on another stream; the next gf_tree_build_dev
 * must come after this call's kernels.  The neighbours do not depend on the
 * gas tree. */
int gf_hydro_setup_from_tree_dev(int n, const double* pos, const double* vel, const double* mass,
                                 const double* u, void* stream);
int gf_hydro_knn_dev(int k, const double* r_init, void* stream);
/* gf_hydro_knn_dev in parts, so that a caller can enqueue other work (the
 * gravity walk on another stream) while the solve runs: _begin enqueues the
 * set-up and `passes` (0..64) pass launches on the stream and returns
 * without waiting for the device; _more enqueues further passes (a caller
 * that records an event between the set-up and the first pass can order
 * another stream's work after the solve's first kernels); each pass reads
 * its work from the device, so passes after convergence find nothing to do.
 * _end waits for them, checks the device's error flags, enqueues further
 * passes until every particle is resolved, and returns the number of pass
 * launches that had work.  Nothing may read the solve's results before
 * _end. */
int gf_hydro_knn_begin_dev(int k, const double* r_init, int passes, void* stream);
int gf_hydro_knn_more_dev(int passes, void* stream);
int gf_hydro_knn_end_dev(void* stream);
int gf_hydro_forces_dev(double eps, double gamma, double alpha, double beta, double* h,
                        double* rho, double* acc, double* dudt, double* dt, double* divv,
                        void* stream);
/* d_k (distance to the k-th neighbour, itself included) of the last
 * device-resident solve, input order, into a device array. */
int gf_hydro_dk_dev(double* dk, void* stream);
/* The neighbour table of the last device-resident solve as host arrays:
 * nb (n, k) int64 input indices, rows unordered; dk (n) (may be NULL).
 * Returns k. */
int gf_hydro_neighbours(int64_t* nb, double* dk);
/* Statistics of the last neighbour solve: out[0] pass launches, [1] groups
 * the fast pass handed to the heap pass (summed over passes), [2] groups
 * handed on to the global-stack pass, [3] gas-tree nodes, [4] gas groups,
 * [5] gas particles, [6] work items of the first pass (merged groups; 0
 * without a caller's radii) (len entries are written, at most 7). */
int gf_hydro_stats(long long* out, int len);
/* Passes each particle of the last solve took (host array, input order; may
 * be NULL) and their total.  Returns the number of pass launches. */
int gf_hydro_passes(int* per_particle, long long* total);

/* ------------------------------------------------ learned smoothing length
 * No counterpart in the reference (new component): a 4 -> 32 -> 32 -> 1 SiLU
 * MLP (GF_MLP_PARAMS = 1249 FP32 parameters: W1[32][4], b1, W2[32][32], b2,
 * W3[32], b3) on 4 BF16 features per gas particle that proposes the first
 * search radius of the neighbour solve; the plain radius iteration still
 * decides the result.  All pointers are device arrays unless marked host. */

/* Number of MLP parameters. */
int gf_mlp_params(void);
/* Density of the smallest gas-tree cell holding >= k particles around each
 * gas particle of the last gf_hydro_setup_dev, and (out_count, may be NULL)
 * that cell's particle count (input order). */
int gf_hydro_cell_density_dev(int k, double* out, double* out_count, void* stream);
/* Features of each gas particle, free of length and time scales:
 * log(d_k prev / r_cell) with r_cell = (3k / (4 pi n_cell))^(1/3),
 * log(cell count / k), div v(prev) * dt, log(dt / dt_Courant(prev));
 * normalised by norm[0..3] (means) and norm[4..7] (standard deviations) ->
 * x (n, 4) BF16; raw (may be NULL) receives the unnormalised FP32 values. */
int gf_mlp_features_dev(int n, int k, const double* dk_prev, const double* cell_rho,
                        const double* cell_count, const double* divv, const double* dtc_prev,
                        float dt, const float* norm, uint16_t* x, float* raw, void* stream);
/* Network output y for n samples. */
int gf_mlp_forward_dev(int n, const uint16_t* x, const float* params, float* out, void* stream);
/* First search radius r = d_k(prev) exp(y + margin). */
int gf_mlp_propose_dev(int n, const uint16_t* x, const float* params, const double* dk_prev,
                       float margin, double* r, void* stream);
/* Gradient (grad: 1249 floats) of the mean loss over the n samples and the
 * summed loss (loss: 1 float); either may be NULL.  tau <= 0: squared error
 * 0.5 (y - target)^2 (summed loss: squared errors); 0 < tau < 1: pinball
 * loss of the tau-quantile, max(tau (t - y), (1 - tau) (y - t)). */
int gf_mlp_grad_dev(int n, const uint16_t* x, const float* target, const float* params,
                    float tau, float* grad, float* loss, void* stream);
/* One Adam step (step counts from 1) on the mean loss of n samples. */
int gf_mlp_train_step_dev(int n, const uint16_t* x, const float* target, float* params, float* m,
                          float* v, int step, float lr, float beta1, float beta2, float eps,
                          float tau, float* loss, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight training loops, harnesses that
 * forbid hipMalloc): gf_mlp_train_step_dev with its partial gradients in a
 * static device array of the library instead of a buffer it allocates.  It
 * calls no allocator and does not create the library's stream; it runs on
 * the caller's stream and current device.  n <= 65536 samples per call (more
 * is refused with an error); the array holds 2.56 MB on each device that runs
 * it until the process ends.  Calls that overlap on one device race. */
int gf_mlp_train_step_noalloc_dev(int n, const uint16_t* x, const float* target, float* params,
                                  float* m, float* v, int step, float lr, float beta1, float beta2,
                                  float eps, float tau, float* loss, void* stream);
/* gf_mlp_grad_dev with host arrays (x as uint16 BF16 bit patterns). */
int gf_mlp_grad_host(int n, const uint16_t* x, const float* target, const float* params,
                     float tau, float* grad, float* loss);

/* ------------------------------------------------------------ gas physics
 * The sub-grid steps around the force path, in FP64 with the reference's
 * operation order.  Device arrays; gas rows are 0 .. ng-1 and, where a gas
 * index array g (int64, NULL = identity) is given, u and no_cool_until are
 * the simulation's arrays over all n_total particles indexed through g, as
 * self.u[g] in galaxy_formation/simulation.py. */

/* The unit constants of galaxy_formation/units.py:17-35 as this library
 * evaluates them (host array of 8): MSUN, MSUN_G, M_H, X_H, RHO_CGS, U_CGS,
 * MYR, SEC_PER_MYR. */
int gf_gas_constants(double* out);
/* cooling.cooling_function (cooling.py:24-32): Lambda(T) in erg cm^3 / s. */
int gf_cooling_function_dev(int n, const double* T, double* lam, void* stream);
/* cooling.cool (cooling.py:35-43): out = max(u - du, u_floor) with du the
 * energy radiated in dt at density rho, 0 where blocked (u8, may be NULL). */
int gf_cool_dev(int n, const double* u, const double* T, const double* rho, double dt,
                const uint8_t* blocked, double u_floor, double* out, void* stream);
/* Simulation._thermodynamics (simulation.py:97-104) for ng gas rows, in place
 * on u: u[g] = max(u[g] + dt dudt, u_floor), then cooling with
 * T = u[g] / u_per_K, blocked where no_cool_until[g] > t (no_cool_until may be
 * NULL).  dudt and rho are per gas row.  Rows whose g is outside
 * [0, n_total) are skipped and counted for gf_gas_errors_dev. */
int gf_thermo_dev(int ng, const int64_t* g, int n_total, double* u, const double* dudt,
                  const double* rho, const double* no_cool_until, double t, double dt,
                  double u_per_K, double u_floor, void* stream);
/* Rows skipped by gf_thermo_dev since the last call (waits; clears). */
int gf_gas_errors_dev(void* stream);
/* star_formation.form_stars (star_formation.py:35-47), first half: the gas
 * rows of dense (rho > rho_sf), cold (u[g] / u_per_K < 3e4 K) gas, ascending,
 * into rows, and p = 1 - exp(-sf_eff dt / t_ff), t_ff = sqrt(3 pi / (32 rho)).
 * Returns the count m (waits). */
int gf_sf_candidates_dev(int ng, const int64_t* g, int n_total, const double* u,
                         const double* rho, double u_per_K, double rho_sf, double sf_eff,
                         double dt, int64_t* rows, double* p, void* stream);
/* Second half (star_formation.py:44): rows[i] for which uniform[i] < p[i],
 * in order, into out.  The m uniform draws come from the caller's generator
 * (numpy's rng.uniform(size=m) in the reference).  Returns the count (waits). */
int gf_sf_select_dev(int m, const int64_t* rows, const double* p, const double* uniform,
                     int64_t* out, void* stream);
/* The k nearest source points of each query point: cKDTree(src).query(q, k)
 * of star_formation.py:59.  nb (n_q, k) gets source input indices ascending in
 * (distance, index), dist (may be NULL) the distances.  1 <= k <= 256.
 * Waits. */
int gf_knn_query_dev(int n_src, const double* src, int n_q, const double* q, int k, int64_t* nb,
                     double* dist, void* stream);
/* Supernova heating (star_formation.py:60-65): for the (n_new, k) targets
 * (particle indices), du = E / k / (mass[t] MSUN MSUN_G) / U_CGS with
 * E = e_per_g new_mass MSUN MSUN_G, added to u in the order np.add.at adds the
 * flattened list; no_cool_until[targets] = until.  Waits for the index check. */
int gf_feedback_heat_dev(int n_new, int k, const int64_t* targets, const double* new_mass,
                         double e_per_g, double until, int n_total, const double* mass, double* u,
                         double* no_cool_until, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GF_H */

// Gas physics beside the force path (gf_gas.hip): radiative cooling and the
// explicit thermodynamics of a sub-step, star-formation candidates, and the
// supernova feedback (a k-nearest query of new stars against the gas, then
// the heating).
#pragma once
#include "gf_tree.h"

// Unit constants of galaxy_formation/units.py:17-35, evaluated as that file
// evaluates them (same operations, same order, FP64).
struct GfGasUnits {
    double msun, msun_g, m_h, x_h, rho_cgs, u_cgs, myr, sec_per_myr;
};
const GfGasUnits& gf_gas_units();

// Scratch of the gas-physics entry points (grown, never shrunk).
struct GfGas {
    GfTree tree;                   // tree of the source points of a k-nearest query
    int* state = nullptr;          // [0] query iterations exhausted, [1] count
    int* flag = nullptr; size_t cap_flag = 0;
    int* scan_tmp = nullptr; size_t cap_scan = 0;
    uint64_t* key = nullptr; uint32_t* val = nullptr; size_t cap_pair = 0;
    uint64_t* key2 = nullptr; uint32_t* val2 = nullptr;
    double* du = nullptr;
    int* sort_tmp = nullptr; size_t cap_sort = 0;
};
void gf_gas_free(GfGas& G);

// Lambda(T) (erg cm^3 / s), cooling.py:24-32.
int gf_cooling_function(int n, const double* T, double* lam, hipStream_t s);
// cooling.py:35-43: out = max(u - du, u_floor), du = 0 where blocked.
int gf_cool(int n, const double* u, const double* T, const double* rho, double dt,
            const uint8_t* blocked, double u_floor, double* out, hipStream_t s);
// simulation.py:97-104 on the gas: u[g] = max(u[g] + dt dudt, u_floor), then
// cooling with T = u[g] / u_per_K and blocked = no_cool_until[g] > t.  u and
// no_cool_until (may be null: nothing blocked) are indexed through g (null =
// identity), dudt and rho by gas row.  Rows with g outside [0, n_total) are
// skipped and counted; gf_gas_errors reports them.  Does not wait.
int gf_thermo_launch(GfGas& G, int ng, const int64_t* g, int n_total, double* u,
                     const double* dudt, const double* rho, const double* no_cool_until, double t,
                     double dt, double u_per_K, double u_floor, hipStream_t s);
// Rows skipped since the last call (waits; clears the count).
int gf_gas_errors(GfGas& G, hipStream_t s);
// star_formation.py:37-43: gas rows (0 .. ng-1, ascending) of dense, cold gas
// and their probabilities p = 1 - exp(-sf_eff dt / t_ff); u is indexed
// through g like gf_thermo_launch, rho by row.  Returns the count (waits).
int gf_sf_candidates(GfGas& G, int ng, const int64_t* g, int n_total, const double* u,
                     const double* rho, double u_per_K, double rho_sf, double sf_eff, double dt,
                     int64_t* rows, double* p, hipStream_t s);
// star_formation.py:44: the rows whose uniform draw is below p, in order.
// Returns the count (waits).
int gf_sf_select(GfGas& G, int m, const int64_t* rows, const double* p, const double* uniform,
                 int64_t* out, hipStream_t s);
// The k nearest of each query point among the source points (input indices,
// ascending in (distance, index)), as the cKDTree query of
// star_formation.py:59.  dist may be null.  Waits for the result.
int gf_knn_query(GfGas& G, int n_src, const double* src, int n_q, const double* q, int k,
                 int64_t* nb, double* dist, hipStream_t s);
// star_formation.py:60-65: u[targets] += du in the order np.add.at applies
// the flattened (star, neighbour) list; no_cool_until[targets] = until.
int gf_feedback_heat(GfGas& G, int n_new, int k, const int64_t* targets, const double* new_mass,
                     double e_per_g, double until, int n_total, const double* mass, double* u,
                     double* no_cool_until, hipStream_t s);

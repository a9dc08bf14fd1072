// The caller's side of the force path on the device (gf_step.hip): the sum of
// the forces and the time-step rule of galaxy_formation/simulation.py:61-72,
// and the per-particle steps of a block (hierarchical) leapfrog with rungs,
// the pattern of simulation.py:114-149.  FP64, the operations in the order
// those lines write them.
#pragma once
#include "gf_common.h"

// acc = acc_grav + acc_sph on the gas rows 0 .. ng-1 (acc_grav elsewhere);
// dt_ind = 0.3 sqrt(eps / max(|acc|, 1e-30)), on the gas rows the smaller of
// that and dt_c; *dt_grav = 0.3 sqrt(eps / max(max |acc|, 1e-30)) (device).
int gf_combine(int n, int ng, double eps, const double* acc_grav, const double* acc_sph,
               const double* dt_c, double* acc, double* dt_ind, double* dt_grav, hipStream_t s);

// Sub-step k: pos += dt vel for all n particles; active[i] = 1 where k is a
// multiple of 2^(rungs-1-rung[i]), else 0; *count += the number active.
int gf_step_drift(int n, double dt, long long k, int rungs, double* pos, const double* vel,
                  const int64_t* rung, uint8_t* active, int* count, hipStream_t s);

// Half kicks of the active particles by h = dt_max 2^-rung / 2: vel += h acc,
// and on the gas rows u = max(u + h dudt, u_floor).  Before the kick, mode 1
// moves the active particles to the rung of their dt_ind (a longer step only
// where sub-step k is also its boundary) and mode 2 moves every particle to
// it (a new ladder); mode 0 keeps the rungs.  The rung of dt is
// min(max(ceil(log2(max(dt_max / dt, 1))), 0), rungs - 1).
int gf_step_kick(int n, int ng, long long k, int rungs, double dt_max, int mode,
                 const uint8_t* active, const double* acc, const double* dudt,
                 const double* dt_ind, int64_t* rung, double* vel, double* u, double u_floor,
                 hipStream_t s);

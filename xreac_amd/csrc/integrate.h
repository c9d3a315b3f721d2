// Integrator launchers (integrate.hip).
#pragma once
#include "common.h"

hipError_t xr_verlet(const FFDev& ff, int ns, const int* offsets, const int* types, const double* dt_fs, double* x,
                     double* v, const double* F, int stage, double* kinetic, hipStream_t s);
hipError_t xr_fire(int ns, const int* offsets, double* x, double* v, const double* F, double* state, int* status,
                   double ftol, int max_iterations, hipStream_t s);
// hist: three (s, t) arrays of n atoms, newest first; order: per system, 0..3.
hipError_t xr_qeq_extrapolate(int ns, int n, const int* offsets, const int* order, double* state, double* hist,
                              hipStream_t s);

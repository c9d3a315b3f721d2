// Host-side launchers shared between translation units of libxreac_hip.
#pragma once
#include "common.h"

struct BinGeom {
  int nb[3];     // bins per direction
  int reach[3];  // bins searched on each side
  double lo[3];  // grid origin in fractional (periodic) or scaled units
  double w[3];   // bin width in the same units
  int pbc[3];
};

// scan.hip
size_t xr_scan_scratch_ints(int n);
hipError_t xr_exclusive_scan(const int* in, int n, int* out, int* scratch, hipStream_t s);

// neighbor.hip
hipError_t xr_coord_bounds(int n, const double* x, CellDev cell, double* bounds6, hipStream_t s);
hipError_t xr_bin_atoms(int n, const double* x, CellDev cell, BinGeom g, int* home, int* bin, int* bin_count,
                        hipStream_t s);
size_t xr_bin_record_bytes();
hipError_t xr_bin_scatter(int n, const double* x, const int* home, const int* bin, const int* bin_start,
                          int* bin_fill, void* recs, unsigned* err, hipStream_t s);
hipError_t xr_neighbor_count(int n, const double* x, CellDev cell, BinGeom g, const int* home, const int* bin,
                             const int* bin_start, const void* binned, double cutoff, int row_cap, int* counts,
                             unsigned* err, hipStream_t s);
hipError_t xr_neighbor_fill(int n, const double* x, CellDev cell, BinGeom g, const int* home, const int* bin,
                            const int* bin_start, const void* binned, double cutoff, const int* row_ptr, int* nbr_j,
                            int* nbr_s, unsigned* err, hipStream_t s);

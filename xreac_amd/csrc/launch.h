// Host-side launchers shared between translation units of libxreac_hip.
#pragma once
#include "common.h"


// scan.hip
size_t xr_scan_scratch_ints(int n);
hipError_t xr_exclusive_scan(const int* in, int n, int* out, int* scratch, hipStream_t s);
// The scan with every pointer clamped to cap; *total receives the unclamped
// total and err the XR_ERR_CAPACITY bit when it exceeds cap; *max_out (when
// not null) the largest count.
hipError_t xr_scan_clamped(const int* in, int n, int* out, int* scratch, int cap, int* total, unsigned* err,
                           int* max_out, hipStream_t s);
// Clamp of existing pointers ptr[0..n] to cap, with the same total and error word.
hipError_t xr_clamp_ptr(int* ptr, int n, int cap, int* total, unsigned* err, hipStream_t s);

// neighbor.hip
hipError_t xr_coord_bounds(const Systems& S, const double* x, double* bounds, hipStream_t s);
// Nonperiodic directions' bin grids from the bounds, in the device system table (no host read-back).
hipError_t xr_bin_grid(SysDev* sys, int ns, const double* bounds, double rcb, unsigned* err, hipStream_t s);
hipError_t xr_bin_atoms(int n, const double* x, const Systems& S, int* home, int* bin, int* bin_count,
                        hipStream_t s);
size_t xr_bin_record_bytes();
hipError_t xr_bin_scatter(int n, const double* x, const Systems& S, const int* home, const int* bin,
                          const int* bin_start, int* bin_fill, void* recs, hipStream_t s);
hipError_t xr_neighbor_count(int n, const double* x, const Systems& S, const int* home, const int* bin,
                             const int* bin_start, const void* binned, double cutoff, int row_cap, int* counts,
                             hipStream_t s);
// Slab row stride (keys per row) of the fill instance for rows up to max_row.
int xr_fill_stride(int max_row);
// Sorted rows to slab[i * stride, i * stride + counts[i]) as (j << 32 | image code).  With a gate (one
// word per system), only the rows of systems whose word is nonzero are written.
hipError_t xr_neighbor_fill(int n, const double* x, const Systems& S, const int* home, const int* bin,
                            const int* bin_start, const void* binned, double cutoff, int* counts, void* slab,
                            int max_row, unsigned* flags, hipStream_t s, const unsigned* gate = nullptr);
// Skin reuse: gate[s] = 1 when system s needs a new skin list (force, an atom moved more than allowance
// from xref, or a type differs from tref); the gated copy of positions and types after a rebuild; and
// the kept skin rows (stride skin_stride) filtered to the cutoff into the slab rows (stride `stride`)
// and counts, as the fill writes them.
hipError_t xr_skin_check(const Systems& S, const double* x, const int* types, const double* xref, const int* tref,
                         double allowance, bool force, unsigned* gate, hipStream_t s);
hipError_t xr_skin_commit(int n, const Systems& S, const double* x, const int* types, double* xref, int* tref,
                          const unsigned* gate, hipStream_t s);
hipError_t xr_skin_filter(int n, const double* x, const Systems& S, const void* skin, int skin_stride,
                          const int* skin_counts, double cutoff, int* counts, void* slab, int stride, unsigned* flags,
                          hipStream_t s);
// Supplied row pointers ptr[0..n] clamped to [0, cap]; a malformed pointer
// array (first entry not 0, negative or decreasing entries) flags the row's
// system with XR_ERR_NEIGHBORS.  *total and the capacity bit as xr_clamp_ptr.
hipError_t xr_clamp_supplied(int* row_ptr, int n, int cap, int* total, unsigned* err, const Systems& S,
                             hipStream_t s);
hipError_t xr_pack_supplied(int n, const Systems& S, const int* row_ptr, const int* in_j, const int* in_s,
                            int* nbr_j, int* nbr_s, hipStream_t s);

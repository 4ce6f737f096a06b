/*
 * helix_io.h -- C entry points of the trajectory and checkpoint files (libhelix.so).
 *
 * Energy trace: the legacy CLI's outputEnergy.txt (src/cli_compatibility.cu:80-98,200-238): one row
 * "<time> <energy>" per recorded step, both printed with 8 significant digits (std::setprecision(8),
 * default float format), the time accumulated by repeated addition of dt, and the energy taken from the
 * diagonal only (src/library/energy_observable.cpp:8-39).  The compiled reference keeps the energy in a
 * single-precision Complex before printing; float32_values = 1 reproduces that rounding.
 *
 * Checkpoint: no counterpart in working form in the reference (its snapshot writer stores raw FP32
 * values without a header and is never read back, src/cli_compatibility.cu:130-145,
 * src/initialize.cu:266-285).  This format is versioned, little-endian, and always FP64:
 *   offset  0  char[8]  magic "HLXCKPT1"
 *           8  uint32   format version (1)
 *          12  uint32   header bytes (96)
 *          16  int32[6] n, n_ado, n_modes, n_baths, order, precision of the run that wrote it
 *          40  uint64   steps taken
 *          48  double   time (sum of the time steps taken)
 *          56  double   dt
 *          64  uint64   hierarchy hash (helix_io_hierarchy_hash of the [n_ado][n_modes] rows)
 *          72  uint64   payload hash (FNV-1a 64 of the payload bytes)
 *          80  uint64   payload bytes (= n_ado * n * n * 16)
 *          88  uint64   problem hash (helix_io_problem_hash of the run's coefficients and operators; 0 = none)
 *          96  payload  [n_ado][n][n] complex128, row-major, interleaved (re, im)
 * Return codes follow include/helix_hip.h (0 ok, < 0 contract error).
 */
#ifndef HELIX_IO_H
#define HELIX_IO_H

#include <stdint.h>

#include "helix_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Writes rows (times[r], energies[r]) for r < rows; -1 when the file cannot be written. */
int helix_io_write_energy_trace(const char* path, const double* times, const double* energies, int64_t rows,
	int float32_values);

/* FNV-1a 64 over the int32 hierarchy rows [n_ado][n_modes] (little-endian bytes). */
uint64_t helix_io_hierarchy_hash(const int32_t* rows, int64_t n_ado, int32_t n_modes);

typedef struct helix_io_checkpoint {
	int32_t n, n_ado, n_modes, n_baths, order, precision;
	uint64_t steps;
	double time;
	double dt;
	uint64_t hierarchy_hash;
	uint64_t problem_hash;
} helix_io_checkpoint;

/* FNV-1a 64 over everything that defines the propagation of a problem besides its hierarchy rows: the sizes,
 * Taylor order and dt, the coefficient tables (down_comm, down_anti, self_coef, xi, xi_counter), the
 * Hamiltonian (diagonal, flips, dense or CSR form) and every coupling (diagonal, flips, CSR form).  A
 * checkpoint written by one problem is refused by a run whose hash differs (other bath, H, couplings or dt). */
uint64_t helix_io_problem_hash(const helix_hip_problem* problem);

/* Writes info and the full FP64 state [n_ado][n][n][2]. */
int helix_io_checkpoint_write(const char* path, const helix_io_checkpoint* info, const double* state);

/* Reads the header into info; when state != NULL also the payload (state_doubles must be at least
 * 2 n_ado n^2).  Every header field and the payload hash are checked; a mismatch returns -1 with the
 * reason in helix_io_last_error(). */
int helix_io_checkpoint_read(const char* path, helix_io_checkpoint* info, double* state, uint64_t state_doubles);

const char* helix_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

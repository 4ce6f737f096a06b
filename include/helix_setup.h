/*
 * helix_setup.h -- C entry points of the FP64 host set-up (libhelix.so).
 *
 * They expose what the reference does inside initialize() (src/initialize.cu:342-352: PSD,
 * temperature dependence, operators, hierarchy, initial state) as a tabulated problem that
 * helix_hip_plan_create() consumes.  Used by the Python front (helix_amd/) and by any binding
 * that cannot link C++; C++ callers use helix::setup::Problem or the public API in <helix/helix.h>.
 * Return codes follow include/helix_hip.h (0 ok, < 0 contract error).
 */
#ifndef HELIX_SETUP_H
#define HELIX_SETUP_H

#include "helix_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct helix_model helix_model;

/* The reference's compiled legacy spin glass generalised to N = 2^spins (src/parameters.cu:11-39):
 * one bath (beta, gamma, zeta, K), coupling V = sum S^z + spin flips of spins 0..8, diagonal H. */
int helix_model_legacy(int n, double beta, double gamma, double zeta, int pade_terms, int tiers, int order, double dt,
	int precision, helix_model** out);

/* Seeded transverse-field spin glass H = sum_{i<j} J_ij sz_i sz_j + h sum_i sx_i (J_ij ~ N(0,1) from
 * SplitMix64(seed)) with one sigma^z Drude-Lorentz bath per spin.  lam/gamma/temperature hold one
 * value per spin; beta = 1/T, zeta = 2 lam / gamma, omega0 = 1.  Tiers 0..max_tier. */
int helix_model_spin_glass(int spins, uint64_t seed, double h, const double* lam, const double* gamma,
	const double* temperature, int pade_terms, int max_tier, int order, double dt, int precision, helix_model** out);

/* One instance of the batched workload: from SplitMix64(seed) draw, in this order, spins = 3 + floor(4u),
 * max_tier = 2 + floor(4u), lambda = 0.05 + 0.45u, gamma = 0.5 + 1.5u, T = 0.5 + 1.5u (every spin's bath
 * gets the same lambda, gamma, T); J_ij from a fresh SplitMix64(seed) stream as in helix_model_spin_glass,
 * h = 1, Pade K = 1.  spins_out / max_tier_out may be NULL. */
int helix_model_spin_glass_random(uint64_t seed, int order, double dt, int precision, helix_model** out,
	int* spins_out, int* max_tier_out);

/* A general sparse system (System::from_sparse, include/helix/types.h:145-153): a Hermitian H and
 * n_couplings Hermitian couplings in CSR form (int32 indices, complex values [nnz][2]), one Drude-Lorentz /
 * Pade bath per coupling given by the reference's Bath fields (inverse temperature beta, damping gamma,
 * coupling strength zeta; lambda = zeta gamma / 2), tiers 0..max_tier.  The couplings are concatenated:
 * v_offsets [n_couplings][n+1] (each relative to its own operator), v_nnz [n_couplings], v_columns and
 * v_values the operators' arrays one after the other.  Runs on the generic CSR stage kernels. */
int helix_model_sparse(int n, int h_nnz, const int32_t* h_offsets, const int32_t* h_columns, const double* h_values,
	int n_couplings, const int32_t* v_nnz, const int32_t* v_offsets, const int32_t* v_columns, const double* v_values,
	const double* beta, const double* gamma, const double* zeta, int pade_terms, int max_tier, int order, double dt,
	int precision, helix_model** out);

/* Selects (enable = 1) or deselects the dense MFMA commutator for H, materialising the model's
 * Hamiltonian as [N][N] complex (N must be 16, 32 or 64 when the plan is created). */
int helix_model_use_dense_hamiltonian(helix_model* model, int enable);

const helix_hip_problem* helix_model_problem(const helix_model* model);
int helix_model_rows(const helix_model* model, int32_t* rows /* [n_ado][n_modes] */);
double helix_model_energy_diagonal(const helix_model* model, const double* rho0);
double helix_model_energy_trace(const helix_model* model, const double* rho0);
void helix_model_free(helix_model* model);

/* building blocks, exposed for known-answer tests */
int helix_setup_pade(int pade_terms, double* poles, double* residues);
int helix_setup_spin_glass_couplings(int spins, uint64_t seed, double* J /* [spins][spins] */);
long helix_setup_hierarchy_size(int modes, int tiers);

const char* helix_setup_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

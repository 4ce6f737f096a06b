// heom_setup.hpp -- FP64 host set-up of a HEOM problem for the gfx950 plan.
//
// Builds what the reference's initialize() builds (src/initialize.cu:125-352): Pade poles/residues,
// bath coefficients, spin operators, the hierarchy and its up/down tables, the per-ADO coefficient
// slots of initSparse() (src/liouville.cu:685-747) -- for any number of baths -- and fills the
// helix_hip_problem descriptor of include/helix_hip.h.
#pragma once

#include "helix_hip.h"

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace helix::setup {

using cdouble = std::complex<double>;

// Pade [K-1/K] decomposition of the Bose function (src/psd/psd.cu:12-113): poles xi_k, residues eta_k.
struct Pade {
	std::vector<double> poles;
	std::vector<double> residues;
};
Pade pade_bose(int K);

// Bath constants of src/initialize.cu:160-218 for one Drude-Lorentz bath.
struct BathConstants {
	std::vector<double> nu;   // nu_0 = gamma, nu_k = xi_k / beta
	cdouble kxi;              // KXi (real)
	cdouble kxi_counter;      // KXiCounter = i zeta gamma / (2 omega0)
	cdouble ktheta_comm;      // KThetaCommutate
	cdouble ktheta_anti;      // KThetaAntiCommutate
	std::vector<cdouble> kphi; // KPhi_k, k = 1..K
};
BathConstants bath_constants(double beta, double gamma, double zeta, double omega0, int K);

// Hierarchy of multi-indices with sum < tiers in the reference's lexicographic order
// (src/initialize_detail.h:159-214) and its up/down tables (:234-269) by combinatorial ranking.
struct Hierarchy {
	int modes = 0;
	int tiers = 0;
	int64_t size = 0;
	std::vector<int32_t> rows;  // [size][modes]
	std::vector<int32_t> up;    // [size][modes], size when absent
	std::vector<int32_t> down;  // [size][modes], size when absent
	std::vector<int32_t> tier;  // [size]
};
int64_t hierarchy_size(int modes, int tiers);
Hierarchy build_hierarchy(int modes, int tiers);

// Spin-structured operator: diag + sum_f c_f X_{mask_f}.
struct SpinOperator {
	std::vector<double> diag;
	std::vector<int32_t> masks;
	std::vector<double> coefs;
};

// Legacy spin-glass operators (src/initialize.cu:95-157, src/initialize_detail.h:12-157).
struct LegacyModel {
	int N = 1024;
	double omega0 = 1.0, dlong = 1.0, dtrans = 1.0, jn_aver = 0.6;
	bool is_square = false;
};
int legacy_spin_count(int N);
SpinOperator legacy_hamiltonian(const LegacyModel& m);
SpinOperator legacy_coupling(const LegacyModel& m);

// Seeded transverse-field spin glass: H = sum_{i<j} J_ij sz_i sz_j + h sum_i sx_i, V_b = sz_b.
std::vector<double> spin_glass_couplings(int spins, uint64_t seed);
SpinOperator spin_glass_hamiltonian(int spins, const std::vector<double>& J, double h);
SpinOperator sigma_z(int spins, int spin);

// SplitMix64 stream with the 53-bit uniform and Box-Muller normal used for every seeded draw.
struct SplitMix64 {
	uint64_t state;
	explicit SplitMix64(uint64_t seed) : state(seed) {}
	uint64_t next();
	double uniform();
	double normal();
};

// A general sparse operator in CSR form (complex values interleaved).
struct CsrOperator {
	int n = 0;
	std::vector<int32_t> offsets;  // [n+1]
	std::vector<int32_t> columns;
	std::vector<double> values;    // [nnz][2]
	helix_hip_csr_operator view() const;
	SpinOperator diagonal_part() const;  // Re O_ii, no flips
};

struct BathSpec {
	double beta = 3.0, gamma = 1.0, zeta = 0.01, omega0 = 1.0;
	int K = 2;
	SpinOperator coupling;
};

// A fully tabulated problem; owns every array the helix_hip_problem points to.
class Problem {
public:
	Problem(SpinOperator hamiltonian, std::vector<BathSpec> baths, int tiers, int order, double dt, int precision);
	Problem(const Problem&) = delete;
	Problem& operator=(const Problem&) = delete;

	const helix_hip_problem& descriptor() const { return desc_; }
	int n() const { return n_; }
	int64_t ado_count() const { return hier_.size; }
	int modes() const { return hier_.modes; }
	const Hierarchy& hierarchy() const { return hier_; }
	const SpinOperator& hamiltonian() const { return ham_; }
	const std::vector<BathSpec>& baths() const { return baths_; }
	// <H> two ways: diagonal-only Re H_ii Re rho_ii (src/library/energy_observable.cpp:8-39) and Re Tr(H rho)
	double energy_diagonal(const double* rho0) const;
	double energy_trace(const double* rho0) const;
	// Selects the dense (MFMA) commutator with H given as [N][N] complex, row-major.  An empty vector
	// materialises the spin-structured Hamiltonian densely; clear_dense_hamiltonian() returns to it.
	void set_dense_hamiltonian(std::vector<cdouble> h = {});
	void clear_dense_hamiltonian();
	bool dense_hamiltonian() const { return !hdense_.empty(); }
	// General sparse operators: a CSR H replaces the spin-structured one, CSR couplings (one per bath)
	// replace the baths' spin-structured couplings.  Either may be null.
	void set_csr_operators(const CsrOperator* hamiltonian, const std::vector<CsrOperator>* couplings);

private:
	int n_ = 0;
	SpinOperator ham_;
	std::vector<BathSpec> baths_;
	Hierarchy hier_;
	std::vector<int32_t> mode_bath_;
	std::vector<double> down_comm_, down_anti_, self_, xi_, xic_;
	std::vector<helix_hip_spin_operator> couplings_;
	std::vector<double> hdense_;  // [N][N][2] when the dense path is selected
	CsrOperator hcsr_;            // n == 0 when H is spin-structured
	std::vector<CsrOperator> vcsr_;
	helix_hip_csr_operator hcsr_view_{};
	std::vector<helix_hip_csr_operator> vcsr_view_;
	helix_hip_problem desc_{};
};

}  // namespace helix::setup

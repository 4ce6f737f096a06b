#pragma once
// Public API of helix_amd: the reference's helix:: surface (include/helix/types.h:11-209 of
// yjmaxpayne/HELIX) kept name for name so callers drop in; additions are appended only
// (enum values at the end, new members after the existing ones, new overloads beside old ones).
//
// Execution goes to the gfx950 HIP plan of include/helix_hip.h.  Support matrix:
//  - Precision::Double  -> plain FP64 form (complex128 state).
//  - Precision::Single  -> FP32 form (complex64 ADOs, tier-0 accumulation in FP64).
//  - Precision::SplitSingle (added) -> FP32 form with two-term FP32 operator values.
//  - SystemKind::LegacySpinGlass via helix::examples::legacy_spin_glass_system(): the reference's
//    compiled model, any Drude-Lorentz/Pade bath values and hierarchy depth.
//  - SystemKind::Sparse executes for any Hermitian H and couplings (one bath per coupling):
//    spin-structured operators (diagonal + bit-flip terms with one coefficient per flip mask) run on
//    the structured kernels, every other CSR operator on the generic sparse kernels.
//  - ResultMode::FinalState and ResultMode::ObservableTrace execute; Trajectory is validation-only.
//  - Backend::CpuReference is unsupported (no CPU execution path ships in the library).

#include <complex>
#include <cstddef>
#include <memory>
#include <string>
#include <vector>

namespace helix {

enum class Backend {
	LegacyCudaSparse,  // accepted: routed to the HIP plan (the reference's only execution route)
	CudaSparse,        // accepted: routed to the HIP plan
	CpuReference,      // unsupported
	Hip                // added: the gfx950 HIP plan
};

enum class Precision {
	Single,
	Double,
	SplitSingle  // added
};

enum class SystemKind {
	Sparse,
	Dense,
	LegacySpinGlass
};

enum class ResultMode {
	FinalState,
	ObservableTrace,
	Trajectory
};

enum class MatrixStorageOrder {
	// Public reduced_density buffers are row-major.
	RowMajor
};

enum class RunStatus {
	NotStarted,
	Success,
	Failed
};

enum class StatusCode {
	InvalidDimension,
	InvalidRowOffsets,
	InvalidColumnValueSize,
	InvalidColumnIndex,
	InvalidCouplingDimension,
	InvalidRuntimeOption,
	UnsupportedPrecision,
	UnsupportedBackend,
	UnsupportedSystemKind,
	UnsupportedExecution,
	UnsupportedBath,
	UnsupportedHierarchy,
	ConcurrentContextUnsupported,
	BathExponentsInvalid,
	DeviceFailure  // added: a HIP runtime or plan error
};

struct Diagnostic {
	StatusCode code;
	std::string message;
};

class Diagnostics {
public:
	Backend backend = Backend::LegacyCudaSparse;
	Precision precision = Precision::Single;
	std::size_t hilbertSize = 0;
	std::size_t hierarchySize = 0;
	std::size_t steps = 0;
	double timeStep = 0.0;
	int integrationOrder = 0;
	RunStatus status = RunStatus::NotStarted;
	std::vector<std::string> warnings;

	void add(StatusCode code, std::string message);
	bool ok() const noexcept;
	bool hasError(StatusCode code) const noexcept;
	std::string message(StatusCode code) const;
	std::string summary() const;
	const std::vector<Diagnostic>& entries() const noexcept;

private:
	std::vector<Diagnostic> entries_;
};

struct ContextOptions {
	Backend backend = Backend::LegacyCudaSparse;
	Precision precision = Precision::Single;
	bool allowConcurrentContexts = false;
	int device = 0;
	int integrationOrder = 4;
	double timeStep = 0.1;
};

struct SparseOperator {
	std::size_t rows = 0;
	std::size_t cols = 0;
	std::vector<std::size_t> rowOffsets;
	std::vector<std::size_t> columnIndices;
	std::vector<std::complex<double>> values;
};

struct Bath {
	enum class Kind {
		DrudeLorentzPade,
		UserExponents
	};

	Kind kind = Kind::DrudeLorentzPade;
	double inverseTemperature = 0.0;
	double damping = 0.0;
	double couplingStrength = 0.0;
	std::size_t padeTerms = 0;
	std::vector<std::complex<double>> residues;
	std::vector<double> frequencies;
	std::vector<std::complex<double>> exponentCoefficients;
	std::vector<std::complex<double>> exponentRates;

	static Bath drude_lorentz_pade();
	static Bath user_exponents(std::vector<std::complex<double>> coefficients,
		std::vector<std::complex<double>> rates);
	Diagnostics validate_supported() const;
	// added: reorganisation energy lambda, cutoff gamma, temperature T (beta = 1/T, couplingStrength = 2 lambda/gamma)
	static Bath drude_lorentz(double lambda, double gamma, double temperature, std::size_t padeTerms = 1);
};

struct HierarchySpec {
	std::size_t maxDepth = 0;          // number of tiers (tiers 0 .. maxDepth-1)
	std::size_t exponentialTerms = 0;  // padeTerms + 1 per bath
	std::vector<Bath> baths;           // one per coupling operator of a Sparse system

	static HierarchySpec compiled_default(Bath bath = Bath::drude_lorentz_pade());
	Diagnostics validate_supported() const;
};

struct SolverOptions {
	std::size_t steps = 0;
	double timeStep = 0.0;
	ResultMode resultMode = ResultMode::FinalState;
};

struct System {
	SystemKind kind = SystemKind::Sparse;
	SparseOperator hamiltonian;
	std::vector<SparseOperator> couplings;
	Diagnostics diagnostics;

	bool valid() const noexcept;
	static System from_sparse(SparseOperator hamiltonian, std::vector<SparseOperator> couplings = {});
};

struct ReducedDensityShape {
	std::size_t count = 0;
	std::size_t rows = 0;
	std::size_t cols = 0;
	MatrixStorageOrder storageOrder = MatrixStorageOrder::RowMajor;
};

struct RunResult {
	std::vector<double> times;
	std::vector<std::complex<double>> reduced_density;
	ReducedDensityShape reduced_density_shape;
	Diagnostics diagnostics;
	// added (ObservableTrace): one value per entry of `times`
	std::vector<double> energy_diagonal;  // sum_i Re H_ii Re rho_ii (the legacy outputEnergy.txt column)
	std::vector<double> energy;           // Re Tr(H rho_S)

	bool ok() const noexcept;
};

class Context {
public:
	explicit Context(ContextOptions options = {});
	// added: a context bound to an arbitrary problem instead of the legacy model
	Context(ContextOptions options, const System& system, const HierarchySpec& hierarchy);
	~Context() noexcept;

	Context(const Context&) = delete;
	Context& operator=(const Context&) = delete;
	Context(Context&&) noexcept;
	Context& operator=(Context&&) noexcept;

	const Diagnostics& diagnostics() const noexcept;
	bool active() const noexcept;
	void run_steps(std::size_t steps);
	std::vector<std::complex<double>> reduced_density() const;
	void destroy() noexcept;

	// added
	std::size_t hilbert_size() const noexcept;
	std::size_t hierarchy_size() const noexcept;
	double time_step() const noexcept;
	int integration_order() const noexcept;
	double energy_diagonal() const;
	double energy() const;
	void run_steps_observed(std::size_t steps, std::vector<std::complex<double>>& rho_trace,
		std::vector<double>& energy_diagonal_trace, std::vector<double>& energy_trace);
	std::size_t steps_taken() const noexcept;
	double elapsed_time() const noexcept;  // the time steps taken, summed one by one as the legacy CLI does
	// FP64 checkpoint of the whole hierarchy (format in helix_io.h); load_checkpoint needs a context of the
	// same problem (Hilbert size, hierarchy and its hash) and restores the step count and time.
	void save_checkpoint(const std::string& path) const;
	void load_checkpoint(const std::string& path);

private:
	class Impl;
	std::unique_ptr<Impl> impl_;
};

class HEOMSolver {
public:
	HEOMSolver();
	explicit HEOMSolver(ContextOptions options);
	HEOMSolver(Context& context, System system, Bath bath, HierarchySpec hierarchy, SolverOptions options = {});

	Diagnostics validate_options() const;
	RunResult run_steps(std::size_t steps);
	RunResult run(const System& system, const HierarchySpec& hierarchy, const SolverOptions& options) const;

private:
	ContextOptions options_;
	Context* context_ = nullptr;
	System system_;
	Bath bath_;
	HierarchySpec hierarchy_;
	SolverOptions solverOptions_;
};

}  // namespace helix

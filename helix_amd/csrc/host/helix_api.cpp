// helix_api.cpp -- the public helix:: API (include/helix/types.h) over the HIP plan's C-ABI.
//
// Follows the reference's src/helix.cpp:18-772 in behaviour: option validation into Diagnostics,
// one active Context per process (:73-104), HEOMSolver converting exceptions into Diagnostics
// (:705-719, :752-769), RunResult shape (:708-713).  The legacy runtime session
// (src/library/legacy_runtime_session.cu:53-99) is replaced by a helix_hip_plan.
#include <helix/helix.h>

#include "heom_setup.hpp"
#include "helix_hip.h"
#include "helix_io.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <sstream>
#include <stdexcept>

namespace helix {

const char* runtimeVersion() noexcept
{
	return versionString();
}

void Diagnostics::add(StatusCode code, std::string message)
{
	entries_.push_back({code, std::move(message)});
	status = RunStatus::Failed;
}

bool Diagnostics::ok() const noexcept
{
	return entries_.empty();
}

bool Diagnostics::hasError(StatusCode code) const noexcept
{
	for(const auto& d : entries_)
		if(d.code == code) return true;
	return false;
}

std::string Diagnostics::message(StatusCode code) const
{
	for(const auto& d : entries_)
		if(d.code == code) return d.message;
	return {};
}

std::string Diagnostics::summary() const
{
	std::ostringstream s;
	for(std::size_t i = 0; i < entries_.size(); ++i) s << (i ? "; " : "") << entries_[i].message;
	return s.str();
}

const std::vector<Diagnostic>& Diagnostics::entries() const noexcept
{
	return entries_;
}

namespace {

constexpr const char* kLifecycleMessage =
	"helix supports a sequential lifecycle: one active Context owns the device session at a time";

std::atomic_bool& active_flag()
{
	static std::atomic_bool flag{false};
	return flag;
}

// One active Context per process (src/helix.cpp:73-104).
class ActiveContextLease {
public:
	ActiveContextLease()
	{
		bool expected = false;
		if(!active_flag().compare_exchange_strong(expected, true)) throw std::logic_error(kLifecycleMessage);
	}
	~ActiveContextLease() noexcept { active_flag().store(false); }
	ActiveContextLease(const ActiveContextLease&) = delete;
	ActiveContextLease& operator=(const ActiveContextLease&) = delete;
};

void append(Diagnostics& to, const Diagnostics& from)
{
	for(const auto& d : from.entries()) to.add(d.code, d.message);
}

void validate_operator(const char* name, const SparseOperator& op, Diagnostics& d)
{
	const std::string n(name);
	if(op.rowOffsets.size() != op.rows + 1) {
		d.add(StatusCode::InvalidRowOffsets, n + ": CSR rowOffsets size must equal rows + 1");
		return;
	}
	if(op.rowOffsets.front() != 0) d.add(StatusCode::InvalidRowOffsets, n + ": CSR rowOffsets must start at zero");
	for(std::size_t i = 1; i < op.rowOffsets.size(); ++i)
		if(op.rowOffsets[i] < op.rowOffsets[i - 1]) {
			d.add(StatusCode::InvalidRowOffsets, n + ": CSR rowOffsets must be monotonically nondecreasing");
			break;
		}
	if(op.columnIndices.size() != op.values.size())
		d.add(StatusCode::InvalidColumnValueSize, n + ": CSR columnIndices and values must have identical sizes");
	if(op.rowOffsets.back() != op.columnIndices.size())
		d.add(StatusCode::InvalidRowOffsets, n + ": CSR final row offset must equal the number of column entries");
	for(const auto c : op.columnIndices)
		if(c >= op.cols) {
			d.add(StatusCode::InvalidColumnIndex, n + ": CSR column index is outside operator dimensions");
			break;
		}
}

Diagnostics validate_system_operators(const SparseOperator& h, const std::vector<SparseOperator>& couplings)
{
	Diagnostics d;
	validate_operator("hamiltonian", h, d);
	if(h.rows != h.cols) d.add(StatusCode::InvalidDimension, "hamiltonian: system operator must be square");
	for(std::size_t i = 0; i < couplings.size(); ++i) {
		const std::string name = "coupling[" + std::to_string(i) + "]";
		validate_operator(name.c_str(), couplings[i], d);
		if(couplings[i].rows != h.rows || couplings[i].cols != h.cols)
			d.add(StatusCode::InvalidCouplingDimension, name + ": coupling operator dimensions must match the system operator");
	}
	return d;
}

Diagnostics validate_context_options(const ContextOptions& o)
{
	Diagnostics d;
	if(o.backend == Backend::CpuReference)
		d.add(StatusCode::UnsupportedBackend, "Backend::CpuReference has no execution path; use Backend::Hip");
	if(o.allowConcurrentContexts) d.add(StatusCode::ConcurrentContextUnsupported, kLifecycleMessage);
	if(o.device < 0) d.add(StatusCode::InvalidRuntimeOption, "device index must be non-negative");
	if(o.integrationOrder <= 0) d.add(StatusCode::InvalidRuntimeOption, "integrationOrder must be positive");
	if(o.integrationOrder > 8) d.add(StatusCode::InvalidRuntimeOption, "integrationOrder above 8 is not supported");
	if(!(o.timeStep > 0.0)) d.add(StatusCode::InvalidRuntimeOption, "timeStep must be positive");
	return d;
}

Diagnostics validate_solver_options(const SolverOptions& o)
{
	Diagnostics d;
	if(o.timeStep < 0.0) d.add(StatusCode::InvalidRuntimeOption, "SolverOptions::timeStep must be non-negative");
	if(o.resultMode == ResultMode::Trajectory)
		d.add(StatusCode::UnsupportedExecution, "ResultMode::Trajectory (every ADO at every step) is not produced");
	return d;
}

int precision_code(Precision p)
{
	switch(p) {
	case Precision::Double: return HELIX_HIP_FP64;
	case Precision::SplitSingle: return HELIX_HIP_SPLIT_FP32;
	default: return HELIX_HIP_FP32;
	}
}

// CSR -> diag + bit flips with one real coefficient per mask, or a reason why not.
bool to_spin_operator(const SparseOperator& op, setup::SpinOperator& out, std::string& why)
{
	const std::size_t n = op.rows;
	if(n < 8 || (n & (n - 1)) != 0) {
		why = "dimension must be a power of two >= 8";
		return false;
	}
	out.diag.assign(n, 0.0);
	std::map<std::size_t, std::pair<double, std::size_t>> flips;  // mask -> (coef, count)
	for(std::size_t i = 0; i < n; ++i)
		for(std::size_t z = op.rowOffsets[i]; z < op.rowOffsets[i + 1]; ++z) {
			const std::size_t j = op.columnIndices[z];
			const std::complex<double> v = op.values[z];
			if(v.imag() != 0.0) {
				why = "complex operator entries are not supported";
				return false;
			}
			if(i == j) {
				out.diag[i] += v.real();
				continue;
			}
			auto it = flips.find(i ^ j);
			if(it == flips.end()) {
				flips[i ^ j] = {v.real(), 1};
			} else {
				if(it->second.first != v.real()) {
					why = "off-diagonal entries of one flip mask must share one coefficient";
					return false;
				}
				++it->second.second;
			}
		}
	out.masks.clear();
	out.coefs.clear();
	for(const auto& f : flips) {
		if(f.second.second != n) {
			why = "every row must hold each flip term";
			return false;
		}
		if(f.second.first == 0.0) continue;
		out.masks.push_back(static_cast<int32_t>(f.first));
		out.coefs.push_back(f.second.first);
	}
	if(out.masks.size() > 16) {
		why = "more than 16 flip terms";
		return false;
	}
	return true;
}

// SparseOperator (the reference's CSR) -> the set-up's CSR; the plan checks Hermiticity
setup::CsrOperator to_setup_csr(const SparseOperator& op, std::size_t n, const std::string& name)
{
	if(op.rows != n || op.cols != n || op.rowOffsets.size() != n + 1 || op.columnIndices.size() != op.values.size() ||
		op.rowOffsets.back() != op.values.size())
		throw std::invalid_argument(name + ": malformed CSR operator");
	if(n < 8 || (n & (n - 1)) != 0) throw std::invalid_argument(name + ": dimension must be a power of two >= 8");
	setup::CsrOperator c;
	c.n = static_cast<int>(n);
	for(std::size_t v : op.rowOffsets) c.offsets.push_back(static_cast<int32_t>(v));
	for(std::size_t v : op.columnIndices) c.columns.push_back(static_cast<int32_t>(v));
	for(const auto& v : op.values) {
		c.values.push_back(v.real());
		c.values.push_back(v.imag());
	}
	return c;
}

setup::BathSpec bath_spec(const Bath& b, setup::SpinOperator coupling)
{
	setup::BathSpec s;
	s.beta = b.inverseTemperature;
	s.gamma = b.damping;
	s.zeta = b.couplingStrength;
	s.omega0 = 1.0;
	s.K = static_cast<int>(b.padeTerms);
	s.coupling = std::move(coupling);
	return s;
}

std::unique_ptr<setup::Problem> make_problem(const ContextOptions& o, const System* system, const HierarchySpec* hier)
{
	HierarchySpec h = hier ? *hier : HierarchySpec::compiled_default();
	if(h.baths.empty()) h.baths.push_back(Bath::drude_lorentz_pade());
	const int order = o.integrationOrder;
	const int prec = precision_code(o.precision);
	if(!system || system->kind == SystemKind::LegacySpinGlass) {
		setup::LegacyModel m;
		return std::make_unique<setup::Problem>(setup::legacy_hamiltonian(m),
			std::vector<setup::BathSpec>{bath_spec(h.baths.front(), setup::legacy_coupling(m))}, static_cast<int>(h.maxDepth),
			order, o.timeStep, prec);
	}
	if(system->kind != SystemKind::Sparse) throw std::invalid_argument("only Sparse and LegacySpinGlass systems execute");
	std::string why;
	// spin-structured operators (diagonal + bit flips) take the structured kernels; any other Hermitian
	// CSR operator runs on the generic CSR kernels
	setup::SpinOperator ham;
	const bool h_csr = !to_spin_operator(system->hamiltonian, ham, why);
	if(system->couplings.empty()) throw std::invalid_argument("a Sparse system needs at least one coupling operator");
	if(h.baths.size() != 1 && h.baths.size() != system->couplings.size())
		throw std::invalid_argument("HierarchySpec::baths must hold one bath, or one bath per coupling");
	std::vector<setup::SpinOperator> vs(system->couplings.size());
	bool v_csr = false;
	for(std::size_t c = 0; c < system->couplings.size(); ++c)
		if(!to_spin_operator(system->couplings[c], vs[c], why)) v_csr = true;
	const std::size_t n = system->hamiltonian.rows;
	std::vector<setup::CsrOperator> vcsr;
	if(v_csr)
		for(std::size_t c = 0; c < system->couplings.size(); ++c) {
			vcsr.push_back(to_setup_csr(system->couplings[c], n, "coupling[" + std::to_string(c) + "]"));
			vs[c] = vcsr.back().diagonal_part();
		}
	setup::CsrOperator hcsr;
	if(h_csr) {
		hcsr = to_setup_csr(system->hamiltonian, n, "hamiltonian");
		ham = hcsr.diagonal_part();
	}
	std::vector<setup::BathSpec> baths;
	for(std::size_t c = 0; c < system->couplings.size(); ++c)
		baths.push_back(bath_spec(h.baths[h.baths.size() == 1 ? 0 : c], std::move(vs[c])));
	auto problem = std::make_unique<setup::Problem>(std::move(ham), std::move(baths), static_cast<int>(h.maxDepth), order,
		o.timeStep, prec);
	if(h_csr || v_csr) problem->set_csr_operators(h_csr ? &hcsr : nullptr, v_csr ? &vcsr : nullptr);
	return problem;
}

}  // namespace

// ------------------------------------------------------------------ Bath

Bath Bath::drude_lorentz_pade()
{
	// compiled defaults of the reference (src/parameters.cu:25-33): beta 3, gamma 1, zeta 0.01, K = 2
	Bath b;
	b.inverseTemperature = 3.0;
	b.damping = 1.0;
	b.couplingStrength = 0.01;
	b.padeTerms = 2;
	return b;
}

Bath Bath::drude_lorentz(double lambda, double gamma, double temperature, std::size_t padeTerms)
{
	Bath b;
	b.inverseTemperature = 1.0 / temperature;
	b.damping = gamma;
	b.couplingStrength = 2.0 * lambda / gamma;
	b.padeTerms = padeTerms;
	return b;
}

Bath Bath::user_exponents(std::vector<std::complex<double>> coefficients, std::vector<std::complex<double>> rates)
{
	Bath b;
	b.kind = Kind::UserExponents;
	b.exponentCoefficients = std::move(coefficients);
	b.exponentRates = std::move(rates);
	return b;
}

Diagnostics Bath::validate_supported() const
{
	Diagnostics d;
	if(kind == Kind::UserExponents) {
		if(exponentCoefficients.empty())
			d.add(StatusCode::BathExponentsInvalid, "User-provided bath exponent coefficients must not be empty");
		if(exponentRates.empty()) d.add(StatusCode::BathExponentsInvalid, "User-provided bath exponent rates must not be empty");
		if(exponentCoefficients.size() != exponentRates.size())
			d.add(StatusCode::BathExponentsInvalid, "User-provided bath exponent coefficients and rates must have equal sizes");
		for(std::size_t i = 0; i < exponentRates.size(); ++i)
			if(exponentRates[i].real() <= 0.0)
				d.add(StatusCode::BathExponentsInvalid,
					"User-provided bath exponent rate " + std::to_string(i) + " must have a positive real part");
		if(d.ok()) d.add(StatusCode::UnsupportedBath, "User-provided bath exponent sets are validation-only");
		return d;
	}
	if(!(inverseTemperature > 0.0)) d.add(StatusCode::UnsupportedBath, "inverse temperature must be positive");
	if(!(damping > 0.0)) d.add(StatusCode::UnsupportedBath, "bath damping must be positive");
	if(!(couplingStrength >= 0.0)) d.add(StatusCode::UnsupportedBath, "bath coupling strength must be non-negative");
	if(padeTerms < 1 || padeTerms > 16) d.add(StatusCode::UnsupportedBath, "Pade term count must be in [1, 16]");
	if(!frequencies.empty() || !residues.empty())
		d.add(StatusCode::UnsupportedBath, "bath frequencies and residues are derived from the Pade decomposition");
	return d;
}

HierarchySpec HierarchySpec::compiled_default(Bath bath)
{
	HierarchySpec h;
	h.maxDepth = 3;  // JMax (src/parameters.cu:36)
	h.exponentialTerms = bath.padeTerms + 1;
	h.baths.push_back(std::move(bath));
	return h;
}

Diagnostics HierarchySpec::validate_supported() const
{
	Diagnostics d;
	if(maxDepth < 1) d.add(StatusCode::UnsupportedHierarchy, "maxDepth (number of tiers) must be at least 1");
	if(baths.empty()) {
		d.add(StatusCode::UnsupportedHierarchy, "the hierarchy needs at least one bath");
		return d;
	}
	for(const auto& b : baths) {
		if(b.kind == Bath::Kind::DrudeLorentzPade && exponentialTerms != b.padeTerms + 1)
			d.add(StatusCode::UnsupportedHierarchy, "exponentialTerms must equal padeTerms + 1 for every bath");
		for(const auto& e : b.validate_supported().entries()) d.add(e.code, "hierarchy bath: " + e.message);
	}
	return d;
}

bool System::valid() const noexcept
{
	return diagnostics.ok();
}

System System::from_sparse(SparseOperator hamiltonian, std::vector<SparseOperator> couplings)
{
	System s;
	s.kind = SystemKind::Sparse;
	s.hamiltonian = std::move(hamiltonian);
	s.couplings = std::move(couplings);
	s.diagnostics = validate_system_operators(s.hamiltonian, s.couplings);
	return s;
}

namespace examples {

System legacy_spin_glass_system()
{
	System s;
	s.kind = SystemKind::LegacySpinGlass;
	return s;
}

System spin_glass_system(int spins, unsigned long long seed, double h)
{
	const std::size_t N = std::size_t(1) << spins;
	auto to_csr = [N](const setup::SpinOperator& op) {
		SparseOperator c;
		c.rows = c.cols = N;
		c.rowOffsets.push_back(0);
		for(std::size_t i = 0; i < N; ++i) {
			std::vector<std::pair<std::size_t, double>> row{{i, op.diag[i]}};
			for(std::size_t f = 0; f < op.masks.size(); ++f) row.push_back({i ^ static_cast<std::size_t>(op.masks[f]), op.coefs[f]});
			std::sort(row.begin(), row.end());
			for(const auto& e : row) {
				c.columnIndices.push_back(e.first);
				c.values.emplace_back(e.second, 0.0);
			}
			c.rowOffsets.push_back(c.columnIndices.size());
		}
		return c;
	};
	const auto J = setup::spin_glass_couplings(spins, seed);
	std::vector<SparseOperator> couplings;
	for(int s = 0; s < spins; ++s) couplings.push_back(to_csr(setup::sigma_z(spins, s)));
	return System::from_sparse(to_csr(setup::spin_glass_hamiltonian(spins, J, h)), std::move(couplings));
}

}  // namespace examples

bool RunResult::ok() const noexcept
{
	return diagnostics.ok();
}

// --------------------------------------------------------------- Context

class Context::Impl {
public:
	Impl(ContextOptions options, const System* system, const HierarchySpec* hierarchy) : options_(options)
	{
		diagnostics_ = validate_context_options(options_);
		if(!diagnostics_.ok()) throw std::invalid_argument(diagnostics_.summary());
		lease_ = std::make_unique<ActiveContextLease>();
		problem_ = make_problem(options_, system, hierarchy);
		const int rc = helix_hip_plan_create(&problem_->descriptor(), options_.device, &plan_);
		if(rc != 0) throw std::runtime_error(std::string("helix_hip_plan_create: ") + helix_hip_last_error());
		const int N = problem_->n();
		std::vector<double> rho0(static_cast<std::size_t>(N) * N * 2, 0.0);
		rho0[2 * (static_cast<std::size_t>(N - 1) * N + (N - 1))] = 1.0;  // |N-1><N-1| (src/initialize.cu:293)
		check(helix_hip_set_initial(plan_, rho0.data()), "helix_hip_set_initial");
		diagnostics_.backend = Backend::Hip;
		diagnostics_.precision = options_.precision;
		diagnostics_.hilbertSize = static_cast<std::size_t>(N);
		diagnostics_.hierarchySize = static_cast<std::size_t>(problem_->ado_count());
		diagnostics_.timeStep = options_.timeStep;
		diagnostics_.integrationOrder = options_.integrationOrder;
	}
	~Impl() noexcept { destroy(); }

	const Diagnostics& diagnostics() const noexcept { return diagnostics_; }
	bool active() const noexcept { return plan_ != nullptr; }

	void run_steps(std::size_t steps)
	{
		require_active();
		check(helix_hip_step(plan_, steps, 0, nullptr, nullptr), "helix_hip_step");
		advance_clock(steps);
	}

	std::size_t steps_taken() const noexcept { return steps_; }
	double elapsed_time() const noexcept { return time_; }

	void save_checkpoint(const std::string& path) const
	{
		require_active();
		const std::size_t n2 = static_cast<std::size_t>(problem_->n()) * problem_->n();
		std::vector<double> state(static_cast<std::size_t>(problem_->ado_count()) * n2 * 2);
		check(helix_hip_get_state(plan_, state.data()), "helix_hip_get_state");
		const helix_io_checkpoint info = checkpoint_info();
		if(helix_io_checkpoint_write(path.c_str(), &info, state.data()) != 0)
			throw std::runtime_error(std::string("save_checkpoint: ") + helix_io_last_error());
	}

	void load_checkpoint(const std::string& path)
	{
		require_active();
		helix_io_checkpoint info{};
		if(helix_io_checkpoint_read(path.c_str(), &info, nullptr, 0) != 0)
			throw std::runtime_error(std::string("load_checkpoint: ") + helix_io_last_error());
		const helix_io_checkpoint mine = checkpoint_info();
		if(info.n != mine.n || info.n_ado != mine.n_ado || info.n_modes != mine.n_modes ||
			info.hierarchy_hash != mine.hierarchy_hash)
			throw std::invalid_argument("load_checkpoint: the checkpoint belongs to another problem (Hilbert size or "
										"hierarchy differs)");
		if(info.order != mine.order || info.dt != mine.dt)
			throw std::invalid_argument("load_checkpoint: the checkpoint was written with another integration order or "
										"time step");
		if(info.problem_hash != mine.problem_hash)
			throw std::invalid_argument("load_checkpoint: the checkpoint belongs to another problem (bath, Hamiltonian "
										"or coupling coefficients differ)");
		std::vector<double> state(static_cast<std::size_t>(info.n_ado) * info.n * info.n * 2);
		if(helix_io_checkpoint_read(path.c_str(), &info, state.data(), state.size()) != 0)
			throw std::runtime_error(std::string("load_checkpoint: ") + helix_io_last_error());
		check(helix_hip_set_state(plan_, state.data()), "helix_hip_set_state");
		steps_ = static_cast<std::size_t>(info.steps);
		time_ = info.time;
	}

	std::vector<std::complex<double>> reduced_density() const
	{
		require_active();
		const std::size_t n2 = static_cast<std::size_t>(problem_->n()) * problem_->n();
		std::vector<std::complex<double>> out(n2);
		check(helix_hip_get_tier0(plan_, reinterpret_cast<double*>(out.data())), "helix_hip_get_tier0");
		return out;
	}

	void run_steps_observed(std::size_t steps, std::vector<std::complex<double>>& rho, std::vector<double>& ediag,
		std::vector<double>& etr)
	{
		require_active();
		const std::size_t n2 = static_cast<std::size_t>(problem_->n()) * problem_->n();
		rho.assign((steps + 1) * n2, {});
		check(helix_hip_get_tier0(plan_, reinterpret_cast<double*>(rho.data())), "helix_hip_get_tier0");
		if(steps)
			check(helix_hip_step(plan_, steps, 1, reinterpret_cast<double*>(rho.data() + n2), nullptr), "helix_hip_step");
		advance_clock(steps);
		ediag.resize(steps + 1);
		etr.resize(steps + 1);
		for(std::size_t s = 0; s <= steps; ++s) {
			const double* r = reinterpret_cast<const double*>(rho.data() + s * n2);
			ediag[s] = problem_->energy_diagonal(r);
			etr[s] = problem_->energy_trace(r);
		}
	}

	double energy_diagonal() const
	{
		auto r = reduced_density();
		return problem_->energy_diagonal(reinterpret_cast<const double*>(r.data()));
	}

	double energy() const
	{
		auto r = reduced_density();
		return problem_->energy_trace(reinterpret_cast<const double*>(r.data()));
	}

	void destroy() noexcept
	{
		if(plan_) helix_hip_plan_destroy(plan_);
		plan_ = nullptr;
		lease_.reset();
	}

	const setup::Problem* problem() const { return problem_.get(); }
	const ContextOptions& options() const { return options_; }

private:
	static void check(int rc, const char* what)
	{
		if(rc != 0) throw std::runtime_error(std::string(what) + ": " + helix_hip_last_error());
	}
	void require_active() const
	{
		if(!plan_) throw std::logic_error("the Context was destroyed");
	}
	// step count and time, the time summed step by step (the legacy CLI's `time += Param::Step`)
	void advance_clock(std::size_t steps)
	{
		for(std::size_t s = 0; s < steps; ++s) time_ += options_.timeStep;
		steps_ += steps;
	}
	helix_io_checkpoint checkpoint_info() const
	{
		helix_io_checkpoint info{};
		info.n = problem_->n();
		info.n_ado = static_cast<int32_t>(problem_->ado_count());
		info.n_modes = problem_->modes();
		info.n_baths = static_cast<int32_t>(problem_->baths().size());
		info.order = options_.integrationOrder;
		info.precision = static_cast<int32_t>(options_.precision);
		info.steps = steps_;
		info.time = time_;
		info.dt = options_.timeStep;
		const auto& rows = problem_->hierarchy().rows;
		info.hierarchy_hash = helix_io_hierarchy_hash(rows.data(), problem_->ado_count(), problem_->modes());
		info.problem_hash = helix_io_problem_hash(&problem_->descriptor());
		return info;
	}

	ContextOptions options_;
	Diagnostics diagnostics_;
	std::unique_ptr<ActiveContextLease> lease_;
	std::unique_ptr<setup::Problem> problem_;
	helix_hip_plan* plan_ = nullptr;
	std::size_t steps_ = 0;
	double time_ = 0.0;
};

Context::Context(ContextOptions options) : impl_(std::make_unique<Impl>(options, nullptr, nullptr)) {}

Context::Context(ContextOptions options, const System& system, const HierarchySpec& hierarchy)
	: impl_(std::make_unique<Impl>(options, &system, &hierarchy))
{
}

Context::~Context() noexcept = default;
Context::Context(Context&&) noexcept = default;
Context& Context::operator=(Context&&) noexcept = default;

const Diagnostics& Context::diagnostics() const noexcept
{
	static const Diagnostics empty;
	return impl_ ? impl_->diagnostics() : empty;
}

bool Context::active() const noexcept
{
	return impl_ && impl_->active();
}

void Context::run_steps(std::size_t steps)
{
	if(!impl_) throw std::logic_error("Cannot run a moved-from Context");
	impl_->run_steps(steps);
}

std::vector<std::complex<double>> Context::reduced_density() const
{
	if(!impl_) throw std::logic_error("Cannot read reduced density from a moved-from Context");
	return impl_->reduced_density();
}

void Context::destroy() noexcept
{
	if(impl_) impl_->destroy();
}

std::size_t Context::hilbert_size() const noexcept
{
	return impl_ ? static_cast<std::size_t>(impl_->problem()->n()) : 0;
}

std::size_t Context::hierarchy_size() const noexcept
{
	return impl_ ? static_cast<std::size_t>(impl_->problem()->ado_count()) : 0;
}

double Context::time_step() const noexcept
{
	return impl_ ? impl_->options().timeStep : 0.0;
}

int Context::integration_order() const noexcept
{
	return impl_ ? impl_->options().integrationOrder : 0;
}

double Context::energy_diagonal() const
{
	if(!impl_) throw std::logic_error("moved-from Context");
	return impl_->energy_diagonal();
}

double Context::energy() const
{
	if(!impl_) throw std::logic_error("moved-from Context");
	return impl_->energy();
}

void Context::run_steps_observed(std::size_t steps, std::vector<std::complex<double>>& rho_trace,
	std::vector<double>& energy_diagonal_trace, std::vector<double>& energy_trace)
{
	if(!impl_) throw std::logic_error("moved-from Context");
	impl_->run_steps_observed(steps, rho_trace, energy_diagonal_trace, energy_trace);
}

std::size_t Context::steps_taken() const noexcept
{
	return impl_ ? impl_->steps_taken() : 0;
}

double Context::elapsed_time() const noexcept
{
	return impl_ ? impl_->elapsed_time() : 0.0;
}

void Context::save_checkpoint(const std::string& path) const
{
	if(!impl_) throw std::logic_error("moved-from Context");
	impl_->save_checkpoint(path);
}

void Context::load_checkpoint(const std::string& path)
{
	if(!impl_) throw std::logic_error("moved-from Context");
	impl_->load_checkpoint(path);
}

// ------------------------------------------------------------ HEOMSolver

HEOMSolver::HEOMSolver() = default;

HEOMSolver::HEOMSolver(ContextOptions options) : options_(options) {}

HEOMSolver::HEOMSolver(Context& context, System system, Bath bath, HierarchySpec hierarchy, SolverOptions options)
	: context_(&context), system_(std::move(system)), bath_(std::move(bath)), hierarchy_(std::move(hierarchy)),
	  solverOptions_(options)
{
}

Diagnostics HEOMSolver::validate_options() const
{
	return validate_context_options(options_);
}

RunResult HEOMSolver::run_steps(std::size_t steps)
{
	RunResult result;
	if(!context_) {
		result.diagnostics.add(StatusCode::InvalidRuntimeOption, "HEOMSolver::run_steps requires a Context-bound solver");
		return result;
	}
	append(result.diagnostics, system_.diagnostics);
	append(result.diagnostics, bath_.validate_supported());
	append(result.diagnostics, hierarchy_.validate_supported());
	append(result.diagnostics, validate_solver_options(solverOptions_));
	if(system_.kind == SystemKind::Dense)
		result.diagnostics.add(StatusCode::UnsupportedSystemKind, "SystemKind::Dense does not execute");
	if(!result.diagnostics.ok()) return result;
	try {
		const std::size_t N = context_->hilbert_size();
		const double dt = context_->time_step();
		if(solverOptions_.resultMode == ResultMode::ObservableTrace) {
			// times as the legacy CLI prints them: summed step by step from the context's clock
			double t = context_->elapsed_time();
			context_->run_steps_observed(steps, result.reduced_density, result.energy_diagonal, result.energy);
			for(std::size_t s = 0; s <= steps; ++s, t += dt) result.times.push_back(t);
			result.reduced_density_shape = {steps + 1, N, N, MatrixStorageOrder::RowMajor};
		} else {
			context_->run_steps(steps);
			result.reduced_density = context_->reduced_density();
			result.reduced_density_shape = {1, N, N, MatrixStorageOrder::RowMajor};
			result.times.push_back(static_cast<double>(steps) * dt);
		}
		result.diagnostics.backend = Backend::Hip;
		result.diagnostics.precision = context_->diagnostics().precision;
		result.diagnostics.hilbertSize = N;
		result.diagnostics.hierarchySize = context_->hierarchy_size();
		result.diagnostics.steps = steps;
		result.diagnostics.timeStep = dt;
		result.diagnostics.integrationOrder = context_->integration_order();
		result.diagnostics.status = RunStatus::Success;
	} catch(const std::exception& e) {
		result.diagnostics.add(StatusCode::DeviceFailure, e.what());
	}
	return result;
}

RunResult HEOMSolver::run(const System& system, const HierarchySpec& hierarchy, const SolverOptions& options) const
{
	RunResult result;
	result.diagnostics = validate_options();
	append(result.diagnostics, system.diagnostics);
	append(result.diagnostics, hierarchy.validate_supported());
	append(result.diagnostics, validate_solver_options(options));
	if(system.kind == SystemKind::Dense) {
		result.diagnostics.add(StatusCode::UnsupportedSystemKind, "SystemKind::Dense does not execute");
		return result;
	}
	if(!result.diagnostics.ok()) return result;
	try {
		ContextOptions co = options_;
		if(options.timeStep > 0.0) co.timeStep = options.timeStep;
		Context context(co, system, hierarchy);
		const Bath bath = hierarchy.baths.empty() ? Bath::drude_lorentz_pade() : hierarchy.baths.front();
		HEOMSolver bound(context, system, bath, hierarchy, options);
		result = bound.run_steps(options.steps);
		context.destroy();
	} catch(const std::invalid_argument& e) {
		result.diagnostics.add(StatusCode::UnsupportedExecution, e.what());
	} catch(const std::exception& e) {
		result.diagnostics.add(StatusCode::InvalidRuntimeOption, e.what());
	}
	return result;
}

}  // namespace helix

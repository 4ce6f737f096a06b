// heom_setup.cpp -- FP64 host set-up of a HEOM problem (see heom_setup.hpp).
#include "heom_setup.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace helix::setup {

// ------------------------------------------------------------------ Pade

namespace {

// Eigenvalues (ascending) of the symmetric tridiagonal matrix with zero diagonal and off-diagonal e,
// by Sturm-sequence bisection to full double precision.
std::vector<double> tridiagonal_eigenvalues(const std::vector<double>& e)
{
	const int n = static_cast<int>(e.size()) + 1;
	double bound = 0.0;
	for(int i = 0; i < n; ++i) {
		const double l = (i > 0) ? std::fabs(e[i - 1]) : 0.0;
		const double r = (i < n - 1) ? std::fabs(e[i]) : 0.0;
		bound = std::max(bound, l + r);
	}
	bound = bound * 1.0001 + 1e-300;
	auto count_below = [&](double x) {
		int c = 0;
		double q = -x;
		if(q < 0) ++c;
		for(int i = 1; i < n; ++i) {
			const double prev = (q == 0.0) ? 1e-300 : q;
			q = -x - e[i - 1] * e[i - 1] / prev;
			if(q < 0) ++c;
		}
		return c;
	};
	std::vector<double> eig(n);
	for(int k = 0; k < n; ++k) {
		double lo = -bound, hi = bound;
		for(int it = 0; it < 200; ++it) {
			const double mid = 0.5 * (lo + hi);
			if(mid <= lo || mid >= hi) break;
			if(count_below(mid) > k)
				hi = mid;
			else
				lo = mid;
		}
		eig[k] = 0.5 * (lo + hi);
	}
	return eig;
}

double b_bose(int m)
{
	return 2.0 * m + 1.0;
}

}  // namespace

Pade pade_bose(int K)
{
	Pade p;
	if(K <= 0) return p;
	// matrix 1: size 2K, off-diagonal 1/sqrt(b(x+1) b(x+2))
	std::vector<double> e1(2 * K - 1);
	for(int x = 0; x < 2 * K - 1; ++x) e1[x] = 1.0 / std::sqrt(b_bose(x + 1) * b_bose(x + 2));
	const std::vector<double> ev1 = tridiagonal_eigenvalues(e1);
	p.poles.resize(K);
	for(int i = 0; i < K; ++i) p.poles[i] = 2.0 / ev1[2 * K - 1 - i];
	// matrix 2: size 2K-1, off-diagonal 1/sqrt(b(x+2) b(x+3))
	std::vector<double> zeta;
	if(K > 1) {
		std::vector<double> e2(2 * K - 2);
		for(int x = 0; x < 2 * K - 2; ++x) e2[x] = 1.0 / std::sqrt(b_bose(x + 2) * b_bose(x + 3));
		const std::vector<double> ev2 = tridiagonal_eigenvalues(e2);
		const int n2 = 2 * K - 1;
		for(int i = 0; i < K - 1; ++i) zeta.push_back(4.0 / (ev2[n2 - 1 - i] * ev2[n2 - 1 - i]));
	}
	std::vector<double> xi(K);
	for(int i = 0; i < K; ++i) xi[i] = p.poles[i] * p.poles[i];
	const double scaling = K * b_bose(K + 1);
	p.residues.resize(K);
	for(int i = 0; i < K; ++i) {
		double t = (i == K - 1) ? 0.5 * scaling : 0.5 * scaling * (zeta[i] - xi[i]) / (xi[K - 1] - xi[i]);
		for(int j = 0; j < K - 1; ++j) {
			if(j == i) continue;
			t *= (zeta[j] - xi[i]) / (xi[j] - xi[i]);
		}
		p.residues[i] = t;
	}
	return p;
}

BathConstants bath_constants(double beta, double gamma, double zeta, double omega0, int K)
{
	const Pade pd = pade_bose(K);
	BathConstants c;
	c.nu.resize(K + 1);
	c.nu[0] = gamma;
	for(int k = 1; k <= K; ++k) c.nu[k] = pd.poles[k - 1] / beta;
	const double bgh = 0.5 * beta * gamma;
	const double bo = beta * omega0;
	double sigma = 1.0 - bgh / std::tan(bgh);
	for(int k = 1; k <= K; ++k) sigma -= 2.0 * pd.residues[k - 1] * gamma * gamma / (c.nu[k] * c.nu[k] - gamma * gamma);
	c.kxi = cdouble(sigma * (zeta / bo), 0.0);
	c.kxi_counter = cdouble(0.0, zeta * bgh / bo);
	c.ktheta_comm = cdouble(0.0, bgh / std::tan(bgh) * zeta / bo);
	c.ktheta_anti = cdouble(bgh * zeta / bo, 0.0);
	c.kphi.resize(K);
	for(int k = 1; k <= K; ++k)
		c.kphi[k - 1] =
			cdouble(0.0, (zeta / bo) * 2.0 * pd.residues[k - 1] * gamma * gamma / (c.nu[k] * c.nu[k] - gamma * gamma));
	return c;
}

// ------------------------------------------------------------- hierarchy

int64_t hierarchy_size(int modes, int tiers)
{
	// C(modes + tiers - 1, modes)
	if(modes <= 0 || tiers <= 0) return 0;
	long double r = 1.0L;
	for(int q = 1; q <= modes; ++q) r = r * (tiers - 1 + q) / q;
	return static_cast<int64_t>(std::llround(r));
}

Hierarchy build_hierarchy(int modes, int tiers)
{
	Hierarchy h;
	h.modes = modes;
	h.tiers = tiers;
	h.size = hierarchy_size(modes, tiers);
	if(h.size <= 0) throw std::invalid_argument("empty hierarchy");
	if(h.size > (int64_t(1) << 30)) throw std::invalid_argument("hierarchy too large");
	// cnt[m][r]: tuples of length m with sum < r
	std::vector<std::vector<int64_t>> cnt(modes + 1, std::vector<int64_t>(tiers + 1, 0));
	for(int m = 0; m <= modes; ++m)
		for(int r = 1; r <= tiers; ++r) cnt[m][r] = (m == 0) ? 1 : hierarchy_size(m, r);
	auto rank = [&](const int32_t* n) {
		int64_t k = 0;
		int s = 0;
		for(int p = 0; p < modes; ++p) {
			for(int v = 0; v < n[p]; ++v) {
				const int rem = tiers - s - v;
				if(rem > 0) k += cnt[modes - p - 1][rem];
			}
			s += n[p];
		}
		return k;
	};
	const size_t H = static_cast<size_t>(h.size);
	h.rows.assign(H * modes, 0);
	h.tier.assign(H, 0);
	// odometer in lexicographic order, mode 0 most significant
	std::vector<int32_t> cur(modes, 0);
	for(size_t r = 0; r < H; ++r) {
		int s = 0;
		for(int p = 0; p < modes; ++p) {
			h.rows[r * modes + p] = cur[p];
			s += cur[p];
		}
		h.tier[r] = s;
		// advance: increment the last position that can grow, zeroing the tail
		for(int p = modes - 1; p >= 0; --p) {
			int prefix = 0;
			for(int q = 0; q < p; ++q) prefix += cur[q];
			if(prefix + cur[p] + 1 < tiers) {
				++cur[p];
				for(int q = p + 1; q < modes; ++q) cur[q] = 0;
				break;
			}
		}
	}
	h.up.assign(H * modes, static_cast<int32_t>(H));
	h.down.assign(H * modes, static_cast<int32_t>(H));
	std::vector<int32_t> tmp(modes);
	for(size_t r = 0; r < H; ++r) {
		const int32_t* n = &h.rows[r * modes];
		for(int k = 0; k < modes; ++k) {
			std::copy(n, n + modes, tmp.begin());
			if(h.tier[r] + 1 < tiers) {
				tmp[k] = n[k] + 1;
				h.up[r * modes + k] = static_cast<int32_t>(rank(tmp.data()));
			}
			if(n[k] > 0) {
				tmp[k] = n[k] - 1;
				h.down[r * modes + k] = static_cast<int32_t>(rank(tmp.data()));
			}
		}
	}
	return h;
}

// ------------------------------------------------------------ operators

int legacy_spin_count(int N)
{
	return static_cast<int>(std::log(static_cast<double>(N)) / std::log(2.0) + 0.5);
}

namespace {

std::vector<double> legacy_couplings(int spins, double value, bool is_square)
{
	// src/initialize_detail.h:52-114: nearest-neighbour chain, or the 9-spin triangular / square lattices
	static const int tri[32][2] = {{0, 1}, {0, 7}, {1, 0}, {1, 7}, {1, 2}, {1, 8}, {2, 1}, {2, 3}, {2, 8}, {3, 2}, {3, 4},
		{3, 5}, {3, 8}, {4, 3}, {4, 5}, {5, 4}, {5, 6}, {5, 3}, {5, 8}, {6, 5}, {6, 7}, {6, 8}, {7, 6}, {7, 0}, {7, 1},
		{7, 8}, {8, 1}, {8, 2}, {8, 3}, {8, 5}, {8, 6}, {8, 7}};
	static const int sq[24][2] = {{0, 1}, {0, 7}, {1, 0}, {1, 2}, {1, 8}, {2, 1}, {2, 3}, {3, 2}, {3, 4}, {3, 8}, {4, 3},
		{4, 5}, {5, 4}, {5, 6}, {5, 8}, {6, 5}, {6, 7}, {7, 6}, {7, 0}, {7, 8}, {8, 1}, {8, 3}, {8, 5}, {8, 7}};
	std::vector<double> c(static_cast<size_t>(spins) * spins, 0.0);
	if(spins != 9) {
		for(int i = 0; i + 1 < spins; ++i) c[i * spins + i + 1] = c[(i + 1) * spins + i] = value;
	} else if(is_square) {
		for(const auto& e : sq) c[e[0] * 9 + e[1]] = value;
	} else {
		for(const auto& e : tri) c[e[0] * 9 + e[1]] = value;
	}
	return c;
}

double projection(int state, int spin, double omega0)
{
	return ((state >> spin) & 1) ? 0.5 * omega0 : -0.5 * omega0;
}

}  // namespace

SpinOperator legacy_hamiltonian(const LegacyModel& m)
{
	const int spins = legacy_spin_count(m.N);
	const std::vector<double> c = legacy_couplings(spins, m.jn_aver, m.is_square);
	SpinOperator op;
	op.diag.resize(m.N);
	for(int s = 0; s < m.N; ++s) {
		double e = 0.0;
		for(int q = 0; q < spins; ++q) e += projection(s, q, m.omega0);
		for(int j = 0; j < spins; ++j)
			for(int k = 0; k < j; ++k) e += c[j * spins + k] * projection(s, j, m.omega0) * projection(s, k, m.omega0);
		op.diag[s] = e;
	}
	return op;
}

SpinOperator legacy_coupling(const LegacyModel& m)
{
	const int spins = legacy_spin_count(m.N);
	SpinOperator op;
	op.diag.resize(m.N);
	for(int s = 0; s < m.N; ++s) {
		double p = 0.0;
		for(int q = 0; q < spins; ++q) p += projection(s, q, m.omega0);
		op.diag[s] = m.dlong * p;
	}
	// bit flips of spins 0..8 (masks 1..256); the spin-9 flip is absent (src/initialize_detail.h:38-45)
	for(int mask = 1; mask <= 256 && mask < m.N; mask <<= 1) {
		op.masks.push_back(mask);
		op.coefs.push_back(m.dtrans);
	}
	return op;
}

uint64_t SplitMix64::next()
{
	uint64_t z = (state += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

double SplitMix64::uniform()
{
	return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0);
}

double SplitMix64::normal()
{
	const double u1 = 1.0 - uniform();
	const double u2 = uniform();
	return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

std::vector<double> spin_glass_couplings(int spins, uint64_t seed)
{
	SplitMix64 rng(seed);
	std::vector<double> J(static_cast<size_t>(spins) * spins, 0.0);
	for(int i = 0; i < spins; ++i)
		for(int j = i + 1; j < spins; ++j) J[i * spins + j] = rng.normal();
	return J;
}

SpinOperator spin_glass_hamiltonian(int spins, const std::vector<double>& J, double h)
{
	const int N = 1 << spins;
	SpinOperator op;
	op.diag.resize(N);
	for(int s = 0; s < N; ++s) {
		double e = 0.0;
		for(int i = 0; i < spins; ++i)
			for(int j = i + 1; j < spins; ++j) {
				const double si = ((s >> i) & 1) ? 1.0 : -1.0;
				const double sj = ((s >> j) & 1) ? 1.0 : -1.0;
				e += J[i * spins + j] * si * sj;
			}
		op.diag[s] = e;
	}
	if(h != 0.0)
		for(int i = 0; i < spins; ++i) {
			op.masks.push_back(1 << i);
			op.coefs.push_back(h);
		}
	return op;
}

SpinOperator sigma_z(int spins, int spin)
{
	SpinOperator op;
	op.diag.resize(size_t(1) << spins);
	for(size_t s = 0; s < op.diag.size(); ++s) op.diag[s] = ((s >> spin) & 1) ? 1.0 : -1.0;
	return op;
}

// ---------------------------------------------------------------- Problem

Problem::Problem(SpinOperator hamiltonian, std::vector<BathSpec> baths, int tiers, int order, double dt, int precision)
	: ham_(std::move(hamiltonian)), baths_(std::move(baths))
{
	n_ = static_cast<int>(ham_.diag.size());
	if(baths_.empty()) throw std::invalid_argument("at least one bath is required");
	std::vector<int> mode_kind;
	std::vector<double> nu;
	std::vector<cdouble> kphi;
	std::vector<BathConstants> consts;
	for(size_t b = 0; b < baths_.size(); ++b) {
		const BathSpec& s = baths_[b];
		if(static_cast<int>(s.coupling.diag.size()) != n_) throw std::invalid_argument("coupling dimension mismatch");
		consts.push_back(bath_constants(s.beta, s.gamma, s.zeta, s.omega0, s.K));
		for(int k = 0; k <= s.K; ++k) {
			mode_bath_.push_back(static_cast<int32_t>(b));
			mode_kind.push_back(k);
			nu.push_back(consts.back().nu[k]);
			kphi.push_back(k > 0 ? consts.back().kphi[k - 1] : cdouble());
		}
	}
	const int M = static_cast<int>(mode_bath_.size());
	hier_ = build_hierarchy(M, tiers);
	const size_t H = static_cast<size_t>(hier_.size);
	down_comm_.assign(H * M * 2, 0.0);
	down_anti_.assign(H * M * 2, 0.0);
	self_.assign(H * 2, 0.0);
	// per-ADO coefficient slots (src/liouville.cu:723-744)
	for(size_t r = 0; r < H; ++r) {
		const int32_t* n = &hier_.rows[r * M];
		double sum = 0.0;
		for(int k = 0; k < M; ++k) {
			const BathConstants& c = consts[mode_bath_[k]];
			if(mode_kind[k] > 0) {
				const double jnud = -nu[k] * n[k];
				down_comm_[(r * M + k) * 2] = kphi[k].real() * jnud;
				down_comm_[(r * M + k) * 2 + 1] = kphi[k].imag() * jnud;
			} else {
				const double ng = -nu[k] * n[k];
				down_comm_[(r * M + k) * 2] = ng * c.ktheta_comm.real();
				down_comm_[(r * M + k) * 2 + 1] = ng * c.ktheta_comm.imag();
				down_anti_[(r * M + k) * 2] = ng * c.ktheta_anti.real();
				down_anti_[(r * M + k) * 2 + 1] = ng * c.ktheta_anti.imag();
			}
			sum += nu[k] * n[k];
		}
		self_[r * 2] = -sum;
	}
	for(const BathConstants& c : consts) {
		xi_.push_back(c.kxi.real());
		xi_.push_back(c.kxi.imag());
		xic_.push_back(c.kxi_counter.real());
		xic_.push_back(c.kxi_counter.imag());
	}
	for(const BathSpec& s : baths_) {
		helix_hip_spin_operator op{};
		op.diag = s.coupling.diag.data();
		op.nflips = static_cast<int32_t>(s.coupling.masks.size());
		op.flip_masks = s.coupling.masks.data();
		op.flip_coefs = s.coupling.coefs.data();
		couplings_.push_back(op);
	}
	desc_.abi_version = HELIX_HIP_ABI_VERSION;
	desc_.n = n_;
	desc_.n_ado = static_cast<int32_t>(H);
	desc_.n_modes = M;
	desc_.n_baths = static_cast<int32_t>(baths_.size());
	desc_.up = hier_.up.data();
	desc_.down = hier_.down.data();
	desc_.mode_bath = mode_bath_.data();
	desc_.tier = hier_.tier.data();
	desc_.down_comm = down_comm_.data();
	desc_.down_anti = down_anti_.data();
	desc_.self_coef = self_.data();
	desc_.xi = xi_.data();
	desc_.xi_counter = xic_.data();
	desc_.hamiltonian.diag = ham_.diag.data();
	desc_.hamiltonian.nflips = static_cast<int32_t>(ham_.masks.size());
	desc_.hamiltonian.flip_masks = ham_.masks.data();
	desc_.hamiltonian.flip_coefs = ham_.coefs.data();
	desc_.hamiltonian_dense = nullptr;
	desc_.couplings = couplings_.data();
	desc_.precision = precision;
	desc_.order = order;
	desc_.dt = dt;
}

void Problem::set_dense_hamiltonian(std::vector<cdouble> h)
{
	const size_t n2 = static_cast<size_t>(n_) * n_;
	if(h.empty() && hcsr_.n > 0) {
		h.assign(n2, cdouble());
		for(int i = 0; i < n_; ++i)
			for(int z = hcsr_.offsets[i]; z < hcsr_.offsets[i + 1]; ++z)
				h[static_cast<size_t>(i) * n_ + hcsr_.columns[z]] +=
					cdouble(hcsr_.values[2 * static_cast<size_t>(z)], hcsr_.values[2 * static_cast<size_t>(z) + 1]);
	}
	if(h.empty()) {
		h.assign(n2, cdouble());
		for(int i = 0; i < n_; ++i) {
			h[static_cast<size_t>(i) * n_ + i] += ham_.diag[i];
			for(size_t f = 0; f < ham_.masks.size(); ++f) h[static_cast<size_t>(i) * n_ + (i ^ ham_.masks[f])] += ham_.coefs[f];
		}
	}
	if(h.size() != n2) throw std::invalid_argument("dense Hamiltonian must hold N*N entries");
	hdense_.resize(2 * n2);
	for(size_t q = 0; q < n2; ++q) {
		hdense_[2 * q] = h[q].real();
		hdense_[2 * q + 1] = h[q].imag();
	}
	desc_.hamiltonian_dense = hdense_.data();
	desc_.hamiltonian_csr = nullptr;
}

void Problem::clear_dense_hamiltonian()
{
	hdense_.clear();
	desc_.hamiltonian_dense = nullptr;
	desc_.hamiltonian_csr = (hcsr_.n > 0) ? &hcsr_view_ : nullptr;
}

helix_hip_csr_operator CsrOperator::view() const
{
	helix_hip_csr_operator v{};
	v.nnz = static_cast<int32_t>(columns.size());
	v.row_offsets = offsets.data();
	v.columns = columns.data();
	v.values = values.data();
	return v;
}

SpinOperator CsrOperator::diagonal_part() const
{
	SpinOperator op;
	op.diag.assign(n, 0.0);
	for(int i = 0; i < n; ++i)
		for(int z = offsets[i]; z < offsets[i + 1]; ++z)
			if(columns[z] == i) op.diag[i] += values[2 * static_cast<size_t>(z)];
	return op;
}

void Problem::set_csr_operators(const CsrOperator* hamiltonian, const std::vector<CsrOperator>* couplings)
{
	if(hamiltonian) {
		if(hamiltonian->n != n_ || static_cast<int>(hamiltonian->offsets.size()) != n_ + 1)
			throw std::invalid_argument("CSR Hamiltonian dimension mismatch");
		hcsr_ = *hamiltonian;
		ham_ = hcsr_.diagonal_part();
		hcsr_view_ = hcsr_.view();
		desc_.hamiltonian.diag = ham_.diag.data();
		desc_.hamiltonian.nflips = 0;
		desc_.hamiltonian.flip_masks = nullptr;
		desc_.hamiltonian.flip_coefs = nullptr;
		desc_.hamiltonian_csr = &hcsr_view_;
	}
	if(couplings) {
		if(couplings->size() != baths_.size()) throw std::invalid_argument("one CSR coupling per bath is required");
		vcsr_ = *couplings;
		vcsr_view_.clear();
		for(size_t b = 0; b < vcsr_.size(); ++b) {
			if(vcsr_[b].n != n_ || static_cast<int>(vcsr_[b].offsets.size()) != n_ + 1)
				throw std::invalid_argument("CSR coupling dimension mismatch");
			baths_[b].coupling = vcsr_[b].diagonal_part();
			couplings_[b].diag = baths_[b].coupling.diag.data();
			couplings_[b].nflips = 0;
			couplings_[b].flip_masks = nullptr;
			couplings_[b].flip_coefs = nullptr;
			vcsr_view_.push_back(vcsr_[b].view());
		}
		desc_.couplings = couplings_.data();
		desc_.couplings_csr = vcsr_view_.data();
	}
}

double Problem::energy_diagonal(const double* rho0) const
{
	double e = 0.0;
	for(int i = 0; i < n_; ++i) e += ham_.diag[i] * rho0[2 * (static_cast<size_t>(i) * n_ + i)];
	return e;
}

double Problem::energy_trace(const double* rho0) const
{
	if(!hdense_.empty()) {
		// Re Tr(H rho) = sum_ij Re(H_ij rho_ji)
		double e = 0.0;
		for(int i = 0; i < n_; ++i)
			for(int j = 0; j < n_; ++j) {
				const size_t hij = 2 * (static_cast<size_t>(i) * n_ + j), rji = 2 * (static_cast<size_t>(j) * n_ + i);
				e += hdense_[hij] * rho0[rji] - hdense_[hij + 1] * rho0[rji + 1];
			}
		return e;
	}
	if(hcsr_.n > 0) {
		double e = 0.0;
		for(int i = 0; i < n_; ++i)
			for(int z = hcsr_.offsets[i]; z < hcsr_.offsets[i + 1]; ++z) {
				const int j = hcsr_.columns[z];
				const size_t rji = 2 * (static_cast<size_t>(j) * n_ + i);
				e += hcsr_.values[2 * static_cast<size_t>(z)] * rho0[rji] - hcsr_.values[2 * static_cast<size_t>(z) + 1] * rho0[rji + 1];
			}
		return e;
	}
	double e = energy_diagonal(rho0);
	for(size_t f = 0; f < ham_.masks.size(); ++f)
		for(int i = 0; i < n_; ++i) {
			const int j = i ^ ham_.masks[f];
			e += ham_.coefs[f] * rho0[2 * (static_cast<size_t>(j) * n_ + i)];
		}
	return e;
}

}  // namespace helix::setup

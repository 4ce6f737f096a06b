// setup_capi.cpp -- C entry points of the host set-up (include/helix_setup.h).
#include "heom_setup.hpp"
#include "helix_setup.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

using namespace helix::setup;

struct helix_model {
	std::unique_ptr<Problem> problem;
};

namespace {
thread_local std::string g_setup_error;

int setup_fail(int code, const std::string& msg)
{
	g_setup_error = msg;
	return code;
}
}  // namespace

extern "C" {

int helix_model_legacy(int n, double beta, double gamma, double zeta, int pade_terms, int tiers, int order, double dt,
	int precision, helix_model** out)
{
	if(!out) return setup_fail(HELIX_HIP_EINVAL, "out is null");
	*out = nullptr;
	try {
		LegacyModel m;
		m.N = n;
		BathSpec bath;
		bath.beta = beta;
		bath.gamma = gamma;
		bath.zeta = zeta;
		bath.K = pade_terms;
		bath.coupling = legacy_coupling(m);
		auto model = std::make_unique<helix_model>();
		model->problem = std::make_unique<Problem>(legacy_hamiltonian(m), std::vector<BathSpec>{bath}, tiers, order, dt,
			precision);
		*out = model.release();
		return 0;
	} catch(const std::exception& e) {
		return setup_fail(HELIX_HIP_EINVAL, e.what());
	}
}

int helix_model_spin_glass(int spins, uint64_t seed, double h, const double* lam, const double* gamma,
	const double* temperature, int pade_terms, int max_tier, int order, double dt, int precision, helix_model** out)
{
	if(!out || !lam || !gamma || !temperature) return setup_fail(HELIX_HIP_EINVAL, "null argument");
	*out = nullptr;
	if(spins < 3 || spins > 12) return setup_fail(HELIX_HIP_EINVAL, "spins must be in [3, 12]");
	try {
		const std::vector<double> J = spin_glass_couplings(spins, seed);
		std::vector<BathSpec> baths;
		for(int s = 0; s < spins; ++s) {
			BathSpec b;
			b.beta = 1.0 / temperature[s];
			b.gamma = gamma[s];
			b.zeta = 2.0 * lam[s] / gamma[s];
			b.omega0 = 1.0;
			b.K = pade_terms;
			b.coupling = sigma_z(spins, s);
			baths.push_back(std::move(b));
		}
		auto model = std::make_unique<helix_model>();
		model->problem = std::make_unique<Problem>(spin_glass_hamiltonian(spins, J, h), std::move(baths), max_tier + 1,
			order, dt, precision);
		*out = model.release();
		return 0;
	} catch(const std::exception& e) {
		return setup_fail(HELIX_HIP_EINVAL, e.what());
	}
}

int helix_model_spin_glass_random(uint64_t seed, int order, double dt, int precision, helix_model** out,
	int* spins_out, int* max_tier_out)
{
	SplitMix64 rng(seed);
	const int spins = 3 + std::min(static_cast<int>(rng.uniform() * 4.0), 3);
	const int max_tier = 2 + std::min(static_cast<int>(rng.uniform() * 4.0), 3);
	const double lam = 0.05 + 0.45 * rng.uniform();
	const double gamma = 0.5 + 1.5 * rng.uniform();
	const double temp = 0.5 + 1.5 * rng.uniform();
	if(spins_out) *spins_out = spins;
	if(max_tier_out) *max_tier_out = max_tier;
	const std::vector<double> l(spins, lam), g(spins, gamma), t(spins, temp);
	return helix_model_spin_glass(spins, seed, 1.0, l.data(), g.data(), t.data(), 1, max_tier, order, dt, precision, out);
}

int helix_model_sparse(int n, int h_nnz, const int32_t* h_offsets, const int32_t* h_columns, const double* h_values,
	int n_couplings, const int32_t* v_nnz, const int32_t* v_offsets, const int32_t* v_columns, const double* v_values,
	const double* beta, const double* gamma, const double* zeta, int pade_terms, int max_tier, int order, double dt,
	int precision, helix_model** out)
{
	if(!out || !h_offsets || !v_nnz || !v_offsets || !beta || !gamma || !zeta || n_couplings <= 0)
		return setup_fail(HELIX_HIP_EINVAL, "null argument");
	*out = nullptr;
	if(n < 8 || (n & (n - 1)) != 0) return setup_fail(HELIX_HIP_EINVAL, "n must be a power of two >= 8");
	try {
		auto csr = [n](int nnz, const int32_t* off, const int32_t* col, const double* val) {
			CsrOperator o;
			o.n = n;
			o.offsets.assign(off, off + n + 1);
			if(o.offsets.front() != 0 || o.offsets.back() != nnz) throw std::invalid_argument("bad CSR row offsets");
			if(nnz > 0) {
				o.columns.assign(col, col + nnz);
				o.values.assign(val, val + 2 * static_cast<size_t>(nnz));
			}
			return o;
		};
		const CsrOperator h = csr(h_nnz, h_offsets, h_columns, h_values);
		std::vector<CsrOperator> v;
		std::vector<BathSpec> baths;
		size_t base = 0;
		for(int c = 0; c < n_couplings; ++c) {
			v.push_back(csr(v_nnz[c], v_offsets + static_cast<size_t>(c) * (n + 1), v_columns + base, v_values + 2 * base));
			base += static_cast<size_t>(v_nnz[c]);
			BathSpec b;
			b.beta = beta[c];
			b.gamma = gamma[c];
			b.zeta = zeta[c];
			b.omega0 = 1.0;
			b.K = pade_terms;
			b.coupling = v.back().diagonal_part();
			baths.push_back(std::move(b));
		}
		auto model = std::make_unique<helix_model>();
		model->problem = std::make_unique<Problem>(h.diagonal_part(), std::move(baths), max_tier + 1, order, dt, precision);
		model->problem->set_csr_operators(&h, &v);
		*out = model.release();
		return 0;
	} catch(const std::exception& e) {
		return setup_fail(HELIX_HIP_EINVAL, e.what());
	}
}

int helix_model_use_dense_hamiltonian(helix_model* model, int enable)
{
	if(!model) return setup_fail(HELIX_HIP_EINVAL, "model is null");
	try {
		if(enable)
			model->problem->set_dense_hamiltonian();
		else
			model->problem->clear_dense_hamiltonian();
		return 0;
	} catch(const std::exception& e) {
		return setup_fail(HELIX_HIP_EINVAL, e.what());
	}
}

const helix_hip_problem* helix_model_problem(const helix_model* model)
{
	return model ? &model->problem->descriptor() : nullptr;
}

int helix_model_rows(const helix_model* model, int32_t* rows)
{
	if(!model || !rows) return setup_fail(HELIX_HIP_EINVAL, "null argument");
	const auto& r = model->problem->hierarchy().rows;
	std::memcpy(rows, r.data(), r.size() * sizeof(int32_t));
	return 0;
}

double helix_model_energy_diagonal(const helix_model* model, const double* rho0)
{
	return (model && rho0) ? model->problem->energy_diagonal(rho0) : 0.0;
}

double helix_model_energy_trace(const helix_model* model, const double* rho0)
{
	return (model && rho0) ? model->problem->energy_trace(rho0) : 0.0;
}

void helix_model_free(helix_model* model)
{
	delete model;
}

int helix_setup_pade(int pade_terms, double* poles, double* residues)
{
	if(pade_terms <= 0 || !poles || !residues) return setup_fail(HELIX_HIP_EINVAL, "bad arguments");
	const Pade p = pade_bose(pade_terms);
	std::memcpy(poles, p.poles.data(), sizeof(double) * pade_terms);
	std::memcpy(residues, p.residues.data(), sizeof(double) * pade_terms);
	return 0;
}

int helix_setup_spin_glass_couplings(int spins, uint64_t seed, double* J)
{
	if(spins <= 0 || !J) return setup_fail(HELIX_HIP_EINVAL, "bad arguments");
	const std::vector<double> v = spin_glass_couplings(spins, seed);
	std::memcpy(J, v.data(), sizeof(double) * v.size());
	return 0;
}

long helix_setup_hierarchy_size(int modes, int tiers)
{
	return static_cast<long>(hierarchy_size(modes, tiers));
}

const char* helix_setup_last_error(void)
{
	return g_setup_error.c_str();
}

}  // extern "C"

// A general sparse system through the public C++ API: helix::System::from_sparse with CSR operators
// read from a whitespace-separated text file, run by HEOMSolver::run, reduced density printed.
//
//   sparse_system <file>
//
// File: "N B tiers steps dt precision" (precision 0 = Double, 1 = Single), then the Hamiltonian and the
// B couplings, each as "nnz", N+1 row offsets, nnz column indices, nnz (re im) pairs, and after each
// coupling its bath "beta gamma zeta padeTerms".  Output: one line "rho i j re im" per element of the
// final reduced density (17 significant digits).
#include <helix/helix.h>

#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

namespace {

helix::SparseOperator read_operator(std::istream& in, std::size_t n)
{
	std::size_t nnz = 0;
	in >> nnz;
	helix::SparseOperator op;
	op.rows = op.cols = n;
	op.rowOffsets.resize(n + 1);
	op.columnIndices.resize(nnz);
	op.values.resize(nnz);
	for(auto& v : op.rowOffsets) in >> v;
	for(auto& v : op.columnIndices) in >> v;
	for(auto& v : op.values) {
		double re = 0, im = 0;
		in >> re >> im;
		v = {re, im};
	}
	return op;
}

}  // namespace

int main(int argc, char** argv)
{
	if(argc != 2) {
		std::cerr << "usage: sparse_system <file>\n";
		return 2;
	}
	std::ifstream in(argv[1]);
	std::size_t n = 0, nb = 0, tiers = 0, steps = 0;
	double dt = 0.0;
	int precision = 0;
	in >> n >> nb >> tiers >> steps >> dt >> precision;
	if(!in || n == 0 || nb == 0) {
		std::cerr << "cannot read " << argv[1] << '\n';
		return 2;
	}
	auto hamiltonian = read_operator(in, n);
	std::vector<helix::SparseOperator> couplings;
	helix::HierarchySpec hierarchy;
	hierarchy.maxDepth = tiers;
	for(std::size_t b = 0; b < nb; ++b) {
		couplings.push_back(read_operator(in, n));
		helix::Bath bath;
		in >> bath.inverseTemperature >> bath.damping >> bath.couplingStrength >> bath.padeTerms;
		hierarchy.exponentialTerms = bath.padeTerms + 1;
		hierarchy.baths.push_back(bath);
	}
	if(!in) {
		std::cerr << "truncated file\n";
		return 2;
	}
	helix::ContextOptions co;
	co.precision = precision == 0 ? helix::Precision::Double : helix::Precision::Single;
	co.timeStep = dt;
	helix::SolverOptions options;
	options.steps = steps;
	const auto system = helix::System::from_sparse(std::move(hamiltonian), std::move(couplings));
	const auto result = helix::HEOMSolver(co).run(system, hierarchy, options);
	if(!result.ok()) {
		std::cerr << result.diagnostics.summary() << '\n';
		return EXIT_FAILURE;
	}
	std::cout << std::setprecision(17);
	for(std::size_t i = 0; i < n; ++i)
		for(std::size_t j = 0; j < n; ++j) {
			const auto v = result.reduced_density[i * n + j];
			std::cout << "rho " << i << ' ' << j << ' ' << v.real() << ' ' << v.imag() << '\n';
		}
	return EXIT_SUCCESS;
}

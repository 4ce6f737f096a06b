// The reference's documented usage program (examples/cpp/legacy_spin_glass.cpp), edited: the step count is
// taken from argv[1] and the trace of rho_S is printed.  The legacy spin glass through
// helix::HEOMSolver().run(system, hierarchy, options).  The unedited program is built from the reference's
// own file by helix_amd/build.py when that file is present.
//
// Derived from HELIX, MIT License: Copyright (c) 2015 Masashi Tsuchimoto and Yoshitaka Tanimura;
// Copyright (c) 2026 Ye Jun; Copyright (c) 2026 HELIX contributors.
#include <helix/helix.h>

#include <cstdlib>
#include <iostream>

int main(int argc, char** argv)
{
	auto system = helix::examples::legacy_spin_glass_system();
	auto bath = helix::Bath::drude_lorentz_pade();
	auto hierarchy = helix::HierarchySpec::compiled_default(bath);

	helix::SolverOptions options;
	options.steps = (argc > 1) ? static_cast<std::size_t>(std::atoll(argv[1])) : 2;

	auto result = helix::HEOMSolver().run(system, hierarchy, options);
	if(!result.ok()) {
		std::cerr << result.diagnostics.summary() << '\n';
		return EXIT_FAILURE;
	}

	const auto& shape = result.reduced_density_shape;
	if(result.times.size() != shape.count || shape.count != 1 || shape.rows == 0 || shape.cols != shape.rows ||
		result.reduced_density.empty()) {
		std::cerr << "unexpected result shape\n";
		return EXIT_FAILURE;
	}
	double trace = 0.0;
	for(std::size_t i = 0; i < shape.rows; ++i) trace += result.reduced_density[i * shape.rows + i].real();

	std::cout << "helix_cpp_example: steps=" << result.diagnostics.steps << " time=" << result.times.front()
			  << " reduced_density_shape=(" << shape.count << "," << shape.rows << "," << shape.cols << ")"
			  << " trace=" << trace << "\n";
	return EXIT_SUCCESS;
}

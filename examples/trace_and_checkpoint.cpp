// Trajectory output and checkpoint/resume through the public C++ API.
//
//   trace_and_checkpoint legacy-trace <steps> <path>
//     the legacy spin glass (default options: FP32 form, dt 0.1) with ResultMode::ObservableTrace,
//     written as the legacy CLI's outputEnergy.txt (one row per step plus the final row)
//   trace_and_checkpoint resume <directory>
//     a 3-spin glass in FP64: 120 steps in one context, saved after 80; a second context loads the
//     checkpoint and runs the last 40; prints whether the two final reduced densities are bit-identical
#include <helix/helix.h>

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

namespace {

int legacy_trace(std::size_t steps, const std::string& path)
{
	helix::Context context;
	auto system = helix::examples::legacy_spin_glass_system();
	auto bath = helix::Bath::drude_lorentz_pade();
	helix::SolverOptions options;
	options.resultMode = helix::ResultMode::ObservableTrace;
	helix::HEOMSolver solver(context, system, bath, helix::HierarchySpec::compiled_default(bath), options);
	const auto result = solver.run_steps(steps);
	if(!result.ok()) {
		std::cerr << result.diagnostics.summary() << '\n';
		return EXIT_FAILURE;
	}
	if(!helix::write_energy_trace(path, result)) {
		std::cerr << "cannot write " << path << '\n';
		return EXIT_FAILURE;
	}
	std::cout << "rows=" << result.times.size() << " last_time=" << result.times.back() << '\n';
	return EXIT_SUCCESS;
}

int resume(const std::string& dir)
{
	helix::ContextOptions co;
	co.precision = helix::Precision::Double;
	co.timeStep = 0.01;
	const auto system = helix::examples::spin_glass_system(3, 0, 1.0);
	helix::HierarchySpec hierarchy;
	hierarchy.maxDepth = 4;
	hierarchy.exponentialTerms = 2;
	for(int s = 0; s < 3; ++s) hierarchy.baths.push_back(helix::Bath::drude_lorentz(0.1, 1.0, 1.0, 1));
	const std::string ck = dir + "/resume.ckpt";

	std::vector<std::complex<double>> straight;
	double t_saved = 0.0;
	{
		helix::Context a(co, system, hierarchy);
		a.run_steps(80);
		a.save_checkpoint(ck);
		t_saved = a.elapsed_time();
		a.run_steps(40);
		straight = a.reduced_density();
		a.destroy();
	}
	helix::Context b(co, system, hierarchy);
	b.load_checkpoint(ck);
	const bool clock_ok = b.steps_taken() == 80 && b.elapsed_time() == t_saved;
	b.run_steps(40);
	const auto resumed = b.reduced_density();
	const bool equal = resumed.size() == straight.size() &&
		std::memcmp(resumed.data(), straight.data(), resumed.size() * sizeof(resumed[0])) == 0;
	std::cout << "clock_ok=" << clock_ok << " resume_equal=" << equal << " steps=" << b.steps_taken()
			  << " time=" << b.elapsed_time() << '\n';
	return (clock_ok && equal) ? EXIT_SUCCESS : EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv)
{
	try {
		if(argc == 4 && std::string(argv[1]) == "legacy-trace")
			return legacy_trace(static_cast<std::size_t>(std::atoll(argv[2])), argv[3]);
		if(argc == 3 && std::string(argv[1]) == "resume") return resume(argv[2]);
	} catch(const std::exception& e) {
		std::cerr << "error: " << e.what() << '\n';
		return EXIT_FAILURE;
	}
	std::cerr << "usage: trace_and_checkpoint legacy-trace <steps> <path> | resume <directory>\n";
	return 2;
}

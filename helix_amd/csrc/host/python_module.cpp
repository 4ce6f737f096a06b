// python_module.cpp -- the `helix` Python module: pybind11 bindings of the public C++ API.
//
// Same module name, class names, fields and methods as the reference's optional binding
// (bindings/python/helix_bindings.cpp:25-199), so Python code written against it runs unchanged; the
// additions of this build's API (Backend::Hip, Precision::SplitSingle, trajectory energies,
// checkpoints, Bath::drude_lorentz, the spin-glass example) are exposed beside them.
//
// The declaration list follows the reference's binding, MIT License: Copyright (c) 2015 Masashi Tsuchimoto and
// Yoshitaka Tanimura; Copyright (c) 2026 Ye Jun; Copyright (c) 2026 HELIX contributors.
#include <helix/helix.h>

#include <pybind11/complex.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <utility>

namespace py = pybind11;

PYBIND11_MODULE(helix, m)
{
	m.doc() = "Python bindings of the HELIX public C++ API (HIP/gfx950 build).";
	m.attr("__version__") = helix::versionString();
	m.def("runtime_version", &helix::runtimeVersion);

	py::enum_<helix::Backend>(m, "Backend")
		.value("LegacyCudaSparse", helix::Backend::LegacyCudaSparse)
		.value("CudaSparse", helix::Backend::CudaSparse)
		.value("CpuReference", helix::Backend::CpuReference)
		.value("Hip", helix::Backend::Hip)
		.export_values();
	py::enum_<helix::Precision>(m, "Precision")
		.value("Single", helix::Precision::Single)
		.value("Double", helix::Precision::Double)
		.value("SplitSingle", helix::Precision::SplitSingle)
		.export_values();
	py::enum_<helix::SystemKind>(m, "SystemKind")
		.value("Sparse", helix::SystemKind::Sparse)
		.value("Dense", helix::SystemKind::Dense)
		.value("LegacySpinGlass", helix::SystemKind::LegacySpinGlass)
		.export_values();
	py::enum_<helix::ResultMode>(m, "ResultMode")
		.value("FinalState", helix::ResultMode::FinalState)
		.value("ObservableTrace", helix::ResultMode::ObservableTrace)
		.value("Trajectory", helix::ResultMode::Trajectory)
		.export_values();
	py::enum_<helix::MatrixStorageOrder>(m, "MatrixStorageOrder")
		.value("RowMajor", helix::MatrixStorageOrder::RowMajor)
		.export_values();
	py::enum_<helix::RunStatus>(m, "RunStatus")
		.value("NotStarted", helix::RunStatus::NotStarted)
		.value("Success", helix::RunStatus::Success)
		.value("Failed", helix::RunStatus::Failed)
		.export_values();
	py::enum_<helix::StatusCode>(m, "StatusCode")
		.value("InvalidDimension", helix::StatusCode::InvalidDimension)
		.value("InvalidRowOffsets", helix::StatusCode::InvalidRowOffsets)
		.value("InvalidColumnValueSize", helix::StatusCode::InvalidColumnValueSize)
		.value("InvalidColumnIndex", helix::StatusCode::InvalidColumnIndex)
		.value("InvalidCouplingDimension", helix::StatusCode::InvalidCouplingDimension)
		.value("InvalidRuntimeOption", helix::StatusCode::InvalidRuntimeOption)
		.value("UnsupportedPrecision", helix::StatusCode::UnsupportedPrecision)
		.value("UnsupportedBackend", helix::StatusCode::UnsupportedBackend)
		.value("UnsupportedSystemKind", helix::StatusCode::UnsupportedSystemKind)
		.value("UnsupportedExecution", helix::StatusCode::UnsupportedExecution)
		.value("UnsupportedBath", helix::StatusCode::UnsupportedBath)
		.value("UnsupportedHierarchy", helix::StatusCode::UnsupportedHierarchy)
		.value("ConcurrentContextUnsupported", helix::StatusCode::ConcurrentContextUnsupported)
		.value("BathExponentsInvalid", helix::StatusCode::BathExponentsInvalid)
		.value("DeviceFailure", helix::StatusCode::DeviceFailure)
		.export_values();
	py::enum_<helix::EnergyValues>(m, "EnergyValues")
		.value("Float32", helix::EnergyValues::Float32)
		.value("Float64", helix::EnergyValues::Float64)
		.export_values();

	py::class_<helix::Diagnostic>(m, "Diagnostic")
		.def_readwrite("code", &helix::Diagnostic::code)
		.def_readwrite("message", &helix::Diagnostic::message);

	py::class_<helix::Diagnostics>(m, "Diagnostics")
		.def(py::init<>())
		.def_readwrite("backend", &helix::Diagnostics::backend)
		.def_readwrite("precision", &helix::Diagnostics::precision)
		.def_readwrite("hilbertSize", &helix::Diagnostics::hilbertSize)
		.def_readwrite("hierarchySize", &helix::Diagnostics::hierarchySize)
		.def_readwrite("steps", &helix::Diagnostics::steps)
		.def_readwrite("timeStep", &helix::Diagnostics::timeStep)
		.def_readwrite("integrationOrder", &helix::Diagnostics::integrationOrder)
		.def_readwrite("status", &helix::Diagnostics::status)
		.def_readwrite("warnings", &helix::Diagnostics::warnings)
		.def("add", &helix::Diagnostics::add)
		.def("ok", &helix::Diagnostics::ok)
		.def("hasError", &helix::Diagnostics::hasError)
		.def("message", &helix::Diagnostics::message)
		.def("summary", &helix::Diagnostics::summary)
		.def_property_readonly("entries", &helix::Diagnostics::entries, py::return_value_policy::reference_internal);

	py::class_<helix::ContextOptions>(m, "ContextOptions")
		.def(py::init<>())
		.def_readwrite("backend", &helix::ContextOptions::backend)
		.def_readwrite("precision", &helix::ContextOptions::precision)
		.def_readwrite("allowConcurrentContexts", &helix::ContextOptions::allowConcurrentContexts)
		.def_readwrite("device", &helix::ContextOptions::device)
		.def_readwrite("integrationOrder", &helix::ContextOptions::integrationOrder)
		.def_readwrite("timeStep", &helix::ContextOptions::timeStep);

	py::class_<helix::SparseOperator>(m, "SparseOperator")
		.def(py::init<>())
		.def_readwrite("rows", &helix::SparseOperator::rows)
		.def_readwrite("cols", &helix::SparseOperator::cols)
		.def_readwrite("rowOffsets", &helix::SparseOperator::rowOffsets)
		.def_readwrite("columnIndices", &helix::SparseOperator::columnIndices)
		.def_readwrite("values", &helix::SparseOperator::values);

	py::class_<helix::Bath>(m, "Bath")
		.def(py::init<>())
		.def_readwrite("inverseTemperature", &helix::Bath::inverseTemperature)
		.def_readwrite("damping", &helix::Bath::damping)
		.def_readwrite("couplingStrength", &helix::Bath::couplingStrength)
		.def_readwrite("padeTerms", &helix::Bath::padeTerms)
		.def_readwrite("residues", &helix::Bath::residues)
		.def_readwrite("frequencies", &helix::Bath::frequencies)
		.def_static("drude_lorentz_pade", &helix::Bath::drude_lorentz_pade)
		.def_static("drude_lorentz", &helix::Bath::drude_lorentz, py::arg("reorganization"), py::arg("damping"),
			py::arg("temperature"), py::arg("padeTerms") = 1)
		.def("validate_supported", &helix::Bath::validate_supported);

	py::class_<helix::HierarchySpec>(m, "HierarchySpec")
		.def(py::init<>())
		.def_readwrite("maxDepth", &helix::HierarchySpec::maxDepth)
		.def_readwrite("exponentialTerms", &helix::HierarchySpec::exponentialTerms)
		.def_readwrite("baths", &helix::HierarchySpec::baths)
		.def_static("compiled_default", &helix::HierarchySpec::compiled_default,
			py::arg("bath") = helix::Bath::drude_lorentz_pade())
		.def("validate_supported", &helix::HierarchySpec::validate_supported);

	py::class_<helix::SolverOptions>(m, "SolverOptions")
		.def(py::init<>())
		.def_readwrite("steps", &helix::SolverOptions::steps)
		.def_readwrite("timeStep", &helix::SolverOptions::timeStep)
		.def_readwrite("resultMode", &helix::SolverOptions::resultMode);

	py::class_<helix::System>(m, "System")
		.def(py::init<>())
		.def_readwrite("kind", &helix::System::kind)
		.def_readwrite("hamiltonian", &helix::System::hamiltonian)
		.def_readwrite("couplings", &helix::System::couplings)
		.def_readwrite("diagnostics", &helix::System::diagnostics)
		.def("valid", &helix::System::valid)
		.def_static("from_sparse", &helix::System::from_sparse, py::arg("hamiltonian"),
			py::arg("couplings") = std::vector<helix::SparseOperator>{});

	py::class_<helix::ReducedDensityShape>(m, "ReducedDensityShape")
		.def(py::init<>())
		.def_readwrite("count", &helix::ReducedDensityShape::count)
		.def_readwrite("rows", &helix::ReducedDensityShape::rows)
		.def_readwrite("cols", &helix::ReducedDensityShape::cols)
		.def_readwrite("storageOrder", &helix::ReducedDensityShape::storageOrder);

	py::class_<helix::RunResult>(m, "RunResult")
		.def(py::init<>())
		.def_readwrite("times", &helix::RunResult::times)
		.def_readwrite("reduced_density", &helix::RunResult::reduced_density)
		.def_readwrite("reduced_density_shape", &helix::RunResult::reduced_density_shape)
		.def_readwrite("diagnostics", &helix::RunResult::diagnostics)
		.def_readwrite("energy_diagonal", &helix::RunResult::energy_diagonal)
		.def_readwrite("energy", &helix::RunResult::energy)
		.def("ok", &helix::RunResult::ok);

	py::class_<helix::Context>(m, "Context")
		.def(py::init<helix::ContextOptions>(), py::arg("options") = helix::ContextOptions{})
		.def(py::init<helix::ContextOptions, const helix::System&, const helix::HierarchySpec&>(), py::arg("options"),
			py::arg("system"), py::arg("hierarchy"))
		.def("diagnostics", &helix::Context::diagnostics, py::return_value_policy::reference_internal)
		.def("active", &helix::Context::active)
		.def("run_steps", &helix::Context::run_steps, py::call_guard<py::gil_scoped_release>())
		.def("reduced_density", &helix::Context::reduced_density)
		.def("destroy", &helix::Context::destroy)
		.def("hilbert_size", &helix::Context::hilbert_size)
		.def("hierarchy_size", &helix::Context::hierarchy_size)
		.def("time_step", &helix::Context::time_step)
		.def("integration_order", &helix::Context::integration_order)
		.def("energy_diagonal", &helix::Context::energy_diagonal)
		.def("energy", &helix::Context::energy)
		.def("steps_taken", &helix::Context::steps_taken)
		.def("elapsed_time", &helix::Context::elapsed_time)
		.def("save_checkpoint", &helix::Context::save_checkpoint)
		.def("load_checkpoint", &helix::Context::load_checkpoint);

	py::class_<helix::HEOMSolver>(m, "HEOMSolver")
		.def(py::init<>())
		.def(py::init<helix::ContextOptions>(), py::arg("options"))
		.def(py::init<helix::Context&, helix::System, helix::Bath, helix::HierarchySpec, helix::SolverOptions>(),
			py::keep_alive<1, 2>(), py::arg("context"), py::arg("system"), py::arg("bath"), py::arg("hierarchy"),
			py::arg("options") = helix::SolverOptions{})
		.def("validate_options", &helix::HEOMSolver::validate_options)
		.def("run_steps", &helix::HEOMSolver::run_steps)
		.def("run", &helix::HEOMSolver::run, py::arg("system"), py::arg("hierarchy"), py::arg("options"));

	m.def("write_energy_trace",
		py::overload_cast<const std::string&, const helix::RunResult&, helix::EnergyValues>(&helix::write_energy_trace),
		py::arg("path"), py::arg("result"), py::arg("values") = helix::EnergyValues::Float32);

	auto ex = m.def_submodule("examples", "Example systems.");
	ex.def("legacy_spin_glass_system", &helix::examples::legacy_spin_glass_system);
	ex.def("spin_glass_system", &helix::examples::spin_glass_system, py::arg("spins"), py::arg("seed"),
		py::arg("h") = 1.0);
}

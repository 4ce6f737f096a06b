"""CPU tests of the `helix` Python module (pybind11 bindings of the public C++ API): it exposes every
name of the reference's binding (bindings/python/helix_bindings.cpp:25-199) with the reference's
defaults, and its validation paths run without a GPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

REFERENCE_NAMES = [
    "Backend", "Precision", "SystemKind", "ResultMode", "MatrixStorageOrder", "RunStatus", "StatusCode", "Diagnostic",
    "Diagnostics", "ContextOptions", "SparseOperator", "Bath", "HierarchySpec", "SolverOptions", "System",
    "ReducedDensityShape", "RunResult", "Context", "HEOMSolver", "runtime_version", "__version__", "examples",
]


@pytest.fixture(scope="module")
def helix():
    import glob
    if not glob.glob(os.path.join(ROOT, "helix_amd", "helix.*.so")):
        from helix_amd import build
        build.build()
    from helix_amd import helix as mod
    return mod


def test_reference_names_exposed(helix):
    for name in REFERENCE_NAMES:
        assert hasattr(helix, name), name
    assert hasattr(helix.examples, "legacy_spin_glass_system")
    for v in ("LegacyCudaSparse", "CudaSparse", "CpuReference"):
        assert hasattr(helix.Backend, v)
    for v in ("Single", "Double"):
        assert hasattr(helix.Precision, v)
    for v in ("FinalState", "ObservableTrace", "Trajectory"):
        assert hasattr(helix.ResultMode, v)
    for m in ("diagnostics", "active", "run_steps", "reduced_density", "destroy"):
        assert hasattr(helix.Context, m)
    for m in ("validate_options", "run_steps", "run"):
        assert hasattr(helix.HEOMSolver, m)


def test_reference_defaults(helix):
    o = helix.ContextOptions()
    assert o.precision == helix.Precision.Single and o.integrationOrder == 4 and o.timeStep == 0.1
    assert not o.allowConcurrentContexts and o.device == 0
    b = helix.Bath.drude_lorentz_pade()
    assert (b.inverseTemperature, b.damping, b.couplingStrength, b.padeTerms) == (3.0, 1.0, 0.01, 2)
    h = helix.HierarchySpec.compiled_default()
    assert h.maxDepth == 3 and h.exponentialTerms == 3 and len(h.baths) == 1
    assert helix.SolverOptions().resultMode == helix.ResultMode.FinalState


def test_validation_without_gpu(helix):
    op = helix.SparseOperator()
    op.rows = op.cols = 8
    op.rowOffsets = [0, 2]  # wrong length
    s = helix.System.from_sparse(op, [])
    assert not s.valid() and s.diagnostics.hasError(helix.StatusCode.InvalidRowOffsets)
    o = helix.ContextOptions()
    o.backend = helix.Backend.CpuReference
    solver = helix.HEOMSolver(o)
    assert solver.validate_options().hasError(helix.StatusCode.UnsupportedBackend)
    opts = helix.SolverOptions()
    opts.steps = 1
    r = solver.run(helix.examples.legacy_spin_glass_system(), helix.HierarchySpec.compiled_default(), opts)
    assert not r.ok() and r.diagnostics.hasError(helix.StatusCode.UnsupportedBackend)
    opts.resultMode = helix.ResultMode.Trajectory
    r = helix.HEOMSolver().run(helix.examples.legacy_spin_glass_system(), helix.HierarchySpec.compiled_default(), opts)
    assert r.diagnostics.hasError(helix.StatusCode.UnsupportedExecution)

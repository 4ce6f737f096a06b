"""GPU tests of the `helix` Python module: the reference's documented drop-in call
HEOMSolver().run(system, hierarchy, options) from Python, and a seeded spin glass through it against the
CPU FP64 oracle at the plain-form bar (1e-12 normwise relative)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def helix():
    from helix_amd import helix as mod
    return mod


def test_legacy_drop_in_call(helix):
    opts = helix.SolverOptions()
    opts.steps = 2
    r = helix.HEOMSolver().run(helix.examples.legacy_spin_glass_system(), helix.HierarchySpec.compiled_default(), opts)
    assert r.ok(), r.diagnostics.summary()
    sh = r.reduced_density_shape
    assert (sh.count, sh.rows, sh.cols) == (1, 1024, 1024) and r.times == [pytest.approx(0.2)]
    rho = np.array(r.reduced_density).reshape(1024, 1024)
    assert abs(np.trace(rho) - 1) <= 1e-5


def test_spin_glass_against_oracle(helix, oracle_mod, tmp_path):
    co = helix.ContextOptions()
    co.precision = helix.Precision.Double
    co.timeStep = 0.01
    h = helix.HierarchySpec()
    h.maxDepth = 5
    h.exponentialTerms = 2
    h.baths = [helix.Bath.drude_lorentz(0.1, 1.0, 1.0, 1) for _ in range(3)]
    opts = helix.SolverOptions()
    opts.steps = 200
    opts.resultMode = helix.ResultMode.ObservableTrace
    r = helix.HEOMSolver(co).run(helix.examples.spin_glass_system(3, 0), h, opts)
    assert r.ok(), r.diagnostics.summary()
    rho = np.array(r.reduced_density).reshape(201, 8, 8)
    p = oracle_mod.spin_glass_problem(3, 0, L=4)
    ref = p.taylor(p.initial(), 0.01, 200)[0]
    assert float(np.max(np.abs(rho[-1] - ref)) / np.max(np.abs(ref))) <= 1e-12
    assert len(r.energy) == 201 and abs(r.energy[-1] - p.energy(p.taylor(p.initial(), 0.01, 200))) <= 1e-12
    path = tmp_path / "e.txt"
    assert helix.write_energy_trace(str(path), r)
    assert len(path.read_text().splitlines()) == 201


def test_context_checkpoint_refuses_another_bath_or_time_step(helix, tmp_path):
    """Context::load_checkpoint (C++ API) refuses a checkpoint written with another bath or time step, whose
    hierarchy is the same, and resumes one of the same problem."""
    def ctx(lam, dt):
        co = helix.ContextOptions()
        co.precision = helix.Precision.Double
        co.timeStep = dt
        h = helix.HierarchySpec()
        h.maxDepth = 3
        h.exponentialTerms = 2
        h.baths = [helix.Bath.drude_lorentz(lam, 1.0, 1.0, 1) for _ in range(3)]
        return helix.Context(co, helix.examples.spin_glass_system(3, 0), h)

    path = str(tmp_path / "c.ckpt")
    a = ctx(0.1, 0.01)
    a.run_steps(5)
    a.save_checkpoint(path)
    a.destroy()
    for lam, dt, what in ((0.2, 0.01, "coupling coefficients differ"), (0.1, 0.02, "time step")):
        b = ctx(lam, dt)
        with pytest.raises(ValueError, match=what):
            b.load_checkpoint(path)
        b.destroy()
    c = ctx(0.1, 0.01)
    c.load_checkpoint(path)
    assert c.steps_taken() == 5
    c.destroy()

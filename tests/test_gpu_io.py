"""GPU tests of trajectory output and checkpoint/resume, through the C++ API and the C-ABI.

- The legacy spin glass run with ResultMode::ObservableTrace and written by helix::write_energy_trace
  reproduces the reference's golden outputEnergy.txt: every time stamp as printed, every energy within
  the golden trace's own tolerance (1e-5 absolute).
- A run resumed from a checkpoint equals the uninterrupted run bit for bit (FP64 and FP32 forms; the
  FP32 form's tier-0 FP64 master is what the checkpoint stores for ADO 0)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "legacy_outputEnergy.txt")
EXE = os.path.join(ROOT, "helix_amd", "helix_trace_checkpoint")


@pytest.fixture(scope="module")
def hx():
    import helix_amd
    return helix_amd


def test_cpp_energy_trace_matches_golden_file(hx, tmp_path):
    out = tmp_path / "outputEnergy.txt"
    r = subprocess.run([EXE, "legacy-trace", "1980", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = [ln.split() for ln in out.read_text().splitlines()]
    ref = [ln.split() for ln in open(GOLDEN).read().splitlines()]
    assert len(got) == len(ref) == 1981
    assert [g[0] for g in got] == [q[0] for q in ref]  # time column identical as printed
    err = max(abs(float(g[1]) - float(q[1])) for g, q in zip(got, ref))
    print(f"max |dE| against the golden trace: {err:.2e}")
    assert err <= 1e-5


def test_cpp_checkpoint_resume_bitwise(hx, tmp_path):
    r = subprocess.run([EXE, "resume", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "clock_ok=1 resume_equal=1" in r.stdout


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plan_checkpoint_resume_bitwise(hx, tmp_path, precision):
    m = hx.Model.spin_glass(4, 1, L=3, dt=0.01, precision=precision)
    a = hx.Plan(m)
    a.set_initial()
    a.step(50)
    path = str(tmp_path / "p.ckpt")
    a.save_checkpoint(path, steps=50, time=0.5)
    a.step(70)
    b = hx.Plan(m)
    steps, t = b.load_checkpoint(path)
    assert (steps, t) == (50, 0.5)
    b.step(70)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.tier0(), b.tier0())


def test_checkpoint_of_another_problem_is_refused(hx, tmp_path):
    a = hx.Plan(hx.Model.spin_glass(3, 1, L=3))
    a.set_initial()
    path = str(tmp_path / "q.ckpt")
    a.save_checkpoint(path, steps=0, time=0.0)
    with pytest.raises(hx.HelixError, match="another problem"):
        hx.Plan(hx.Model.spin_glass(3, 1, L=2)).load_checkpoint(path)
    # same hierarchy, another bath or time step
    with pytest.raises(hx.HelixError, match="bath, Hamiltonian"):
        hx.Plan(hx.Model.spin_glass(3, 1, L=3, lam=0.2)).load_checkpoint(path)
    with pytest.raises(hx.HelixError, match="time step"):
        hx.Plan(hx.Model.spin_glass(3, 1, L=3, dt=0.02)).load_checkpoint(path)
    assert hx.Plan(hx.Model.spin_glass(3, 1, L=3)).load_checkpoint(path) == (0, 0.0)

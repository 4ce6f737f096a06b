"""helix_hip_step_noalloc (include/helix_hip.h) against helix_hip_step, through the C-ABI.

The allocation-free entry point must give the same states and trace records as helix_hip_step, bit for bit
(both run the same stage kernels on the same plan state; helix_hip_step replays captured 16-step chunks,
whose states equal eager launches bit for bit).  Two calls in a row on one GPU share its device scratch, the
second with a smaller input; records that overflow the 32 MiB scratch are flushed and the next ones reuse it.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _run(hx, model, steps, every, noalloc, record_rho=True, record_diag=True):
    plan = hx.Plan(model)
    plan.set_initial()
    nrec = steps // every if every else 0
    rho = np.zeros((nrec, model.N, model.N), np.complex128) if record_rho else None
    diag = np.zeros((nrec, model.N)) if record_diag else None
    fn = hx.capi.hip().helix_hip_step_noalloc if noalloc else hx.capi.hip().helix_hip_step
    rc = fn(plan._p, steps, every, _pd(rho.view(np.float64)) if rho is not None else None, _pd(diag))
    assert rc == 0, hx.capi.hip().helix_hip_last_error().decode()
    state = plan.get_state()
    t0 = plan.tier0()
    plan.close()
    return rho, diag, state, t0


def _same(a, b):
    for x, y in zip(a, b):
        if x is None or y is None:
            assert x is None and y is None
        else:
            assert np.array_equal(x, y)


def test_noalloc_equals_step_two_calls_second_smaller():
    import helix_amd as hx
    print("noalloc: 4 spins L=3 (165 ADOs), 48 steps / 3, then 3 spins L=2 (28 ADOs), 10 steps / 2", flush=True)
    big = hx.Model.spin_glass(4, 4, L=3, dt=0.01)
    small = hx.Model.spin_glass(3, 0, L=2, dt=0.01)
    a1 = _run(hx, big, 48, 3, noalloc=True)
    a2 = _run(hx, small, 10, 2, noalloc=True)
    _same(a1, _run(hx, big, 48, 3, noalloc=False))
    _same(a2, _run(hx, small, 10, 2, noalloc=False))
    assert a1[0].shape == (16, 16, 16) and a2[0].shape == (5, 8, 8)
    assert abs(np.trace(a2[3]) - 1.0) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_flushes_full_scratch(precision):
    import helix_amd as hx
    # N = 64: one record is 2 * 64^2 + 64 doubles, so 600 records overflow the 32 MiB scratch (508 fit)
    print(f"noalloc: 6 spins L=2 (91 ADOs) {precision}, 1200 steps / 2 (600 records)", flush=True)
    m = hx.Model.spin_glass(6, 1, L=2, dt=0.005, precision=precision)
    _same(_run(hx, m, 1200, 2, noalloc=True), _run(hx, m, 1200, 2, noalloc=False))
    # diagonal only, and no records at all
    _same(_run(hx, m, 40, 1, noalloc=True, record_rho=False), _run(hx, m, 40, 1, noalloc=False, record_rho=False))
    _same(_run(hx, m, 33, 0, noalloc=True, record_rho=False, record_diag=False),
          _run(hx, m, 33, 0, noalloc=False, record_rho=False, record_diag=False))


def test_noalloc_legacy_one_record_per_flush():
    import helix_amd as hx
    # N = 1024: one record (16 MiB + 8 KiB) fills the scratch, so every record is flushed on its own
    print("noalloc: legacy N=1024 (10 ADOs), 3 steps / 1", flush=True)
    m = hx.Model.legacy(N=1024)
    _same(_run(hx, m, 3, 1, noalloc=True), _run(hx, m, 3, 1, noalloc=False))


def test_noalloc_rejects_a_record_beyond_the_scratch():
    import helix_amd as hx
    m = hx.Model.legacy(N=2048, tiers=2)
    plan = hx.Plan(m)
    plan.set_initial()
    rho = np.zeros((1, 2048, 2048), np.complex128)
    rc = hx.capi.hip().helix_hip_step_noalloc(plan._p, 1, 1, _pd(rho.view(np.float64)), None)
    assert rc == -1  # HELIX_HIP_EINVAL, before anything is launched
    assert "staging buffer" in hx.capi.hip().helix_hip_last_error().decode()
    # the diagonal alone fits
    diag = np.zeros((1, 2048))
    assert hx.capi.hip().helix_hip_step_noalloc(plan._p, 1, 1, None, _pd(diag)) == 0
    assert abs(diag.sum() - 1.0) < 1e-12
    plan.close()

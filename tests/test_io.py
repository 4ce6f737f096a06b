"""CPU tests of the trajectory and checkpoint files (include/helix_io.h, libhelix.so).

The energy-trace writer is pinned byte for byte by the reference's own golden file
(examples/outputEnergy.txt): writing its parsed (time, energy) values back with the single-precision
energy rounding of the compiled reference must reproduce the file exactly.  The checkpoint format is
pinned by exact round trips and by the rejection of every corruption its header and hash guard."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "legacy_outputEnergy.txt")


@pytest.fixture(scope="module")
def capi():
    from helix_amd import capi
    if not os.path.exists(capi.HOST_LIB):
        from helix_amd import build
        build.build()
    return capi


def test_energy_trace_writer_reproduces_golden_file(capi, tmp_path):
    rows = [ln.split() for ln in open(GOLDEN).read().splitlines()]
    t = np.array([float(a) for a, _ in rows])
    e = np.array([float(b) for _, b in rows])
    out = tmp_path / "outputEnergy.txt"
    capi.write_energy_trace(str(out), t, e, float32=True)
    assert out.read_bytes() == open(GOLDEN, "rb").read()


def test_energy_trace_times_accumulate_like_the_cli(capi, tmp_path):
    # the CLI sums dt step by step (time += Param::Step); 0.1 summed three times prints as 0.3
    t = np.cumsum([0.0] + [0.1] * 3)
    out = tmp_path / "e.txt"
    capi.write_energy_trace(str(out), t, [6.35, 1.0, -2.5, 1e-9], float32=False)
    assert out.read_text().splitlines() == ["0 6.35", "0.1 1", "0.2 -2.5", "0.3 1e-09"]


def _info(capi, H=7, N=8, M=4):
    return capi.CheckpointInfo(n=N, n_ado=H, n_modes=M, n_baths=2, order=4, precision=0, steps=123,
                               time=1.2300000000000002, dt=0.01, hierarchy_hash=0xDEADBEEF12345678,
                               problem_hash=0x0123456789ABCDEF)


def test_checkpoint_round_trip_exact(capi, tmp_path):
    rng = np.random.default_rng(0)
    info = _info(capi)
    state = rng.standard_normal((7, 8, 8)) + 1j * rng.standard_normal((7, 8, 8))
    path = str(tmp_path / "a.ckpt")
    capi.checkpoint_write(path, info, state)
    got_info, got = capi.checkpoint_read(path)
    assert np.array_equal(got, state)
    for f, _ in capi.CheckpointInfo._fields_:
        assert getattr(got_info, f) == getattr(info, f), f
    assert os.path.getsize(path) == 96 + 7 * 64 * 16


def test_checkpoint_rejects_corruption(capi, tmp_path):
    info = _info(capi)
    state = np.ones((7, 8, 8), np.complex128)
    path = tmp_path / "b.ckpt"
    capi.checkpoint_write(str(path), info, state)
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad.ckpt"
    flipped = bytearray(raw)
    flipped[96 + 100] ^= 1
    bad.write_bytes(bytes(flipped))
    with pytest.raises(capi.HelixError, match="hash"):
        capi.checkpoint_read(str(bad))
    bad.write_bytes(bytes(raw[:-8]))
    with pytest.raises(capi.HelixError, match="truncated"):
        capi.checkpoint_read(str(bad))
    magic = bytearray(raw)
    magic[0] = ord("X")
    bad.write_bytes(bytes(magic))
    with pytest.raises(capi.HelixError, match="magic"):
        capi.checkpoint_read(str(bad))
    size = bytearray(raw)
    size[16] = 9  # n = 9: payload size no longer matches
    bad.write_bytes(bytes(size))
    with pytest.raises(capi.HelixError, match="payload size"):
        capi.checkpoint_read(str(bad))


def test_hierarchy_hash_tells_hierarchies_apart(capi):
    from oracle import oracle as O
    a = O.hierarchy_rows(6, 5)
    b = O.hierarchy_rows(6, 4)
    c = a.copy()
    c[[3, 4]] = c[[4, 3]]  # same rows, different order
    assert capi.hierarchy_hash(a) == capi.hierarchy_hash(a.copy())
    assert len({capi.hierarchy_hash(a), capi.hierarchy_hash(b), capi.hierarchy_hash(c)}) == 3


def test_problem_hash_tells_problems_apart():
    import helix_amd as hx
    base = dict(spins=3, seed=1, lam=0.1, gamma=1.0, T=1.0, L=3, dt=0.01)
    h0 = hx.Model.spin_glass(**base).problem_hash()
    assert h0 != 0 and h0 == hx.Model.spin_glass(**base).problem_hash()
    others = [dict(base, lam=0.2), dict(base, gamma=1.1), dict(base, T=0.9), dict(base, seed=2), dict(base, dt=0.02),
              dict(base, h=0.9)]
    hashes = {hx.Model.spin_glass(**o).problem_hash() for o in others}
    assert len(hashes) == len(others) and h0 not in hashes
    assert hx.Model.spin_glass(**base, order=3).problem_hash() != h0
    # the storage precision does not enter: an FP64 checkpoint resumes an FP32 run of the same problem
    assert hx.Model.spin_glass(**base, precision="fp32").problem_hash() == h0

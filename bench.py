#!/usr/bin/env python3
"""ADO-steps/s of the HEOM propagator on one MI355X.

Workload (default): the plain FP64 transverse-field spin glass with 6 spins (N = 64), one sigma^z
Drude-Lorentz bath per spin (lambda 0.1, gamma 1, T 1, Pade K = 1: M = 12 modes), tiers 0..5
(C(17,5) = 6188 ADOs of 64x64 complex128), Taylor-4 steps of dt = 0.005, synthetic seeded couplings
(seed 1).  One "step" advances every ADO by one full 4-stage step.

value = ADO-steps/s = n_ado * steps / wall time of the timed steps (state resident in HBM).
roofline: the fused stage kernel.  The device stores each Hermitian ADO as P = N(N+1)/2 elements, so the
algorithmic bytes per ADO-step are 14 P s (s = bytes per complex; 14 blocks per step as counted for the
reference's scheme), divided over the m stage launches of a step, against the launch durations measured
with HIP events on the plan's stream (roofline.basis).  The full-block figure 14 N^2 s is reported beside
it as an effective rate only.  cpu_baseline: the FP64 C oracle (OpenMP) timed on this host for one step
of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (spins, seed, max_tier, dt, lam, gamma, T, K)
    "spin_glass_n64_L5": (6, 1, 5, 0.005, 0.1, 1.0, 1.0, 1),
    "spin_glass_n8_L4": (3, 0, 4, 0.01, 0.1, 1.0, 1.0, 1),
    "spin_glass_n256_L3": (8, 2, 3, 0.005, 0.1, 1.0, 1.0, 1),
    "spin_glass_n16_L2": (4, 4, 2, 0.01, 0.1, 1.0, 1.0, 1),
}


def _cpu_cores():
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (the CPU share a GPU gets on its host), else every
    core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, target_s: float = 10.0, max_steps: int = 1000):
    """Taylor-4 steps of the same workload on the FP64 C oracle (OpenMP over ADOs): one step, then, when that took
    less than half of target_s, as many more as fill about target_s (at most max_steps in all), so that a small
    workload's sample is not a few milliseconds."""
    from oracle import oracle as O
    spins, seed, L, dt, lam, gamma, T, K = spec
    cores = _cpu_cores()
    p = O.spin_glass_problem(spins, seed, lam=lam, gamma=gamma, T=T, K=K, L=L, threads=cores)
    r = p.initial()
    t0 = time.perf_counter()
    r = p.taylor(r, dt, 1)
    sec = time.perf_counter() - t0
    steps = 1
    if sec < 0.5 * target_s:
        more = min(max_steps - 1, max(1, int(target_s / max(sec, 1e-6))))
        t0 = time.perf_counter()
        p.taylor(r, dt, more)
        sec += time.perf_counter() - t0
        steps += more
    return {"value": p.H * steps / sec, "unit": "ADO-steps/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{steps} Taylor-4 step(s) of the full workload ({p.H} ADOs, N={p.N}) in {sec:.2f} s"}


def _profiled_traffic(workload, precision, layout):
    """Mean HBM bytes per stage launch from the committed rocprofv3 counter summary, when it covers this
    workload, precision and storage layout (tile kernel)."""
    path = os.path.join(ROOT, "profiles", "stage_traffic_spin_glass_n64_L5.json")
    if workload != "spin_glass_n64_L5" or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if not d.get("layout", "full").startswith(layout):
        return None
    t = d.get("traffic", {}).get(precision)
    return None if t is None else t["mean_bytes_per_launch"]


def _time_single(model_kwargs, steps, warmup=10):
    """ADO-steps/s of `steps` timed steps after `warmup`, and rho_S after all warmup + steps steps."""
    import helix_amd as hx
    m = hx.Model.spin_glass(**model_kwargs)
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(warmup)
    t0 = time.perf_counter()
    plan.step(steps)
    sec = time.perf_counter() - t0
    rho_s = plan.tier0()
    plan.close()
    return m.H * steps / sec, rho_s


def _time_batch(precision, count, steps, warmup=2):
    """ADO-steps/s of the batched workload, its info, and every instance's rho_S after warmup + steps steps."""
    import helix_amd as hx
    models = [hx.Model.spin_glass_random(1000 + i, dt=0.01, precision=precision) for i in range(count)]
    b = hx.Batch(models, graph=True, autotune=True)
    b.set_initial()
    b.step(warmup)
    t0 = time.perf_counter()
    b.step(steps)
    sec = time.perf_counter() - t0
    info = b.info()
    rho_s = [b.tier0(i) for i in range(count)]
    b.close()
    return info["total_ado"] * steps / sec, info, rho_s


def _max_abs(a, b):
    import numpy as np
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def measure_forms(args, fp64_value):
    """The chosen forms beside the plain FP64 line: complex64 and split-FP32 on the same workload, and the
    512-instance batched workload in FP64 and complex64.  Each form reports its ADO-steps/s, the ratio to plain
    FP64 on the same workload, and max|delta rho_S| against the FP64 run of the same workload from the same
    initial state after the same number of steps (warmup + timed)."""
    spins, seed, L, dt, lam, gamma, T, K = WORKLOADS[args.workload]
    kw = dict(spins=spins, seed=seed, lam=lam, gamma=gamma, T=T, K=K, L=L, dt=dt)
    forms = {}
    _, ref = _time_single(dict(kw, precision="fp64"), args.form_steps)
    for prec in ("fp32", "split_fp32"):
        v, rho_s = _time_single(dict(kw, precision=prec), args.form_steps)
        forms[prec] = {"value": v, "x_plain_fp64": v / fp64_value, "max_abs_drho_S": _max_abs(rho_s, ref),
                       "steps_compared": 10 + args.form_steps}
    v64, info, r64 = _time_batch("fp64", 512, args.batch_steps)
    v32, _, r32 = _time_batch("fp32", 512, args.batch_steps)
    seq = _time_one_at_a_time(512, 20)
    forms["batched_512"] = {"n_ado": info["total_ado"], "fp64": v64, "complex64": v32,
                            "launches_per_step_fp64": info["launches_per_step"],
                            "plain_fp64_one_plan_at_a_time": seq,
                            "complex64_x_plain_fp64_same_problems": v32 / seq, "fp64_x_plain_fp64_same_problems": v64 / seq,
                            "complex64_x_batched_fp64": v32 / v64, "complex64_x_plain_fp64_n64": v32 / fp64_value,
                            "complex64_max_abs_drho_S": max(_max_abs(a, b) for a, b in zip(r32, r64)),
                            "steps_compared": 2 + args.batch_steps}
    summary = learned_summary()
    if summary is not None:
        forms["learned_closure_n16_L2"] = summary
    print(f"# forms: {json.dumps(forms)}", file=sys.stderr, flush=True)
    return forms


def _time_one_at_a_time(count, steps):
    """ADO-steps/s of the plain FP64 form on the batched workload's problems, one plan at a time (each plan
    created, warmed up, then `steps` steps timed; creation untimed)."""
    import helix_amd as hx
    t_all, ado_steps = 0.0, 0
    for i in range(count):
        m = hx.Model.spin_glass_random(1000 + i, dt=0.01, precision="fp64")
        p = hx.Plan(m)
        p.set_initial()
        p.step(2)
        t0 = time.perf_counter()
        p.step(steps)
        t_all += time.perf_counter() - t0
        ado_steps += m.H * steps
        p.close()
        m.close()
    return ado_steps / t_all


def copy_probe(nbytes: int) -> int:
    """One hipMemcpy device-to-device of nbytes: a plain copy kernel of known size, so rocprof's
    FETCH_SIZE / WRITE_SIZE for it give the counter scale (bytes read = bytes written = nbytes)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def learned_rows(mlp, seeds, steps=2000, every=20, depths=(6, 5, 4, 3, 2)):
    """Per held-out parameter set: max|d rho_S| over the trajectory against plain FP64 at tiers 0..6 and wall time of
    plain FP64 at every depth and of the tier-2 solver with plain truncation, the physics terminator and the learned
    correction (complex64)."""
    import numpy as np

    import helix_amd as hx
    from helix_amd import closure as cl
    rows = []
    for seed in seeds:
        p = cl.Params.draw(seed)
        row = {"seed": seed}
        ref = None
        for L in depths:  # tiers 0..6 first: the reference
            plan = hx.Plan(p.model(L, "fp64"))
            plan.set_initial()
            plan.step(10)
            plan.set_initial()
            t0 = time.perf_counter()
            tr, _ = plan.step(steps, every, record_rho=True)
            row[f"t_plain_fp64_L{L}_s"] = time.perf_counter() - t0
            plan.close()
            if ref is None:
                ref = tr
            row[f"err_plain_fp64_L{L}"] = float(np.max(np.abs(tr - ref)))
        s = cl.ClosureSolver(p, mlp, keep_tier=2, precision="fp32")
        for mode in ("plain", "physics", "learned"):
            s.use(mode)
            s.run(10)
            t0 = time.perf_counter()
            _, tr = s.run(steps, every)
            row[f"t_{mode}_s"] = time.perf_counter() - t0
            row[f"err_{mode}"] = float(np.max(np.abs(tr - ref)))
        row["n_ado_kept"] = s.model.H
        s.close()
        # shallowest plain FP64 depth at least as accurate as the learned closure (6 = the reference itself)
        row["matched_plain_L"] = min(L for L in sorted(depths) if row[f"err_plain_fp64_L{L}"] <= row["err_learned"])
        rows.append(row)
    return rows


NETWORK = os.path.join(ROOT, "helix_amd", "data", "closure_n16_L2.npz")


def learned_summary(seeds=(9000, 9001, 9002, 9003, 9004, 9005)):
    """Mean errors and time to solution of the learned closure over the held-out sets (None without the network)."""
    import numpy as np

    from helix_amd import closure as cl
    if not os.path.exists(NETWORK):
        return None
    mlp = cl.load_network(NETWORK)
    rows = learned_rows(mlp, seeds)
    mlp.close()
    t_l = sum(r["t_learned_s"] for r in rows)
    return {"kept_tiers": "0..2", "reference": "plain FP64 tiers 0..6", "held_out_seeds": list(seeds),
            "mean_err": {k: float(np.mean([r[f"err_{k}"] for r in rows])) for k in ("plain", "physics", "learned")},
            "learned_better_than_physics": int(sum(r["err_learned"] < r["err_physics"] for r in rows)),
            "matched_plain_L": [r["matched_plain_L"] for r in rows],
            "time_to_solution_x_vs_deep_plain_fp64": sum(r["t_plain_fp64_L6_s"] for r in rows) / t_l,
            "time_to_solution_x_vs_plain_fp64_at_matched_error":
                sum(r[f"t_plain_fp64_L{r['matched_plain_L']}_s"] for r in rows) / t_l}


def bench_learned(args):
    """Learned-closure form: 4 spins (N = 16), tiers 0..2 kept, the physics terminator with its learned correction
    (the committed network helix_amd/data/closure_n16_L2.npz, trained on 4096 parameter sets by
    helix_amd/train_closure.py), complex64, dt 0.01, 2000 steps, on held-out parameter sets (seeds 9000..9005)
    against plain FP64 at tiers 0..6 (3003 ADOs).  value = ADO-steps/s of the learned solver; beside it the error
    (max|d rho_S| over the trajectory) of plain truncation, the physics terminator and the learned correction,
    and time to solution against the deep plain solver and against plain FP64 at the shallowest depth whose
    error is at most the learned closure's."""
    import numpy as np

    from helix_amd import closure as cl
    path = NETWORK
    if not os.path.exists(path):
        print(json.dumps({"error": f"trained network {path} not found (python -m helix_amd.train_closure)"}))
        return 2
    mlp = cl.load_network(path)
    if args.closure_only:  # the learned solver alone (seed 9000, --steps steps), e.g. under a profiler
        s = cl.ClosureSolver(cl.Params.draw(9000), mlp, keep_tier=2, precision="fp32")
        s.use("learned")
        s.run(10)
        t0 = time.perf_counter()
        s.run(args.steps)
        sec = time.perf_counter() - t0
        print(json.dumps({"workload": "learned_closure_n16_L2", "seed": 9000, "steps": args.steps,
                          "ms_per_step": sec * 1e3 / args.steps, "ado_steps_per_s": s.model.H * args.steps / sec}))
        s.close()
        mlp.close()
        return 0
    steps, seeds = 2000, list(range(9000, 9006))
    rows = learned_rows(mlp, seeds, steps)
    for row in rows:
        print(f"# seed {row['seed']}: plain {row['err_plain']:.3e} physics {row['err_physics']:.3e} "
              f"learned {row['err_learned']:.3e} matched plain L={row['matched_plain_L']}", file=sys.stderr, flush=True)
    mlp.close()
    t_learned = sum(r["t_learned_s"] for r in rows)
    t_deep = sum(r["t_plain_fp64_L6_s"] for r in rows)
    t_matched = sum(r[f"t_plain_fp64_L{r['matched_plain_L']}_s"] for r in rows)
    n_kept = rows[0]["n_ado_kept"]
    with np.load(path, allow_pickle=False) as f:
        train_sets = int(f["train_sets"])
    line = {
        "metric": "ADO-steps/sec on 1 MI355X", "value": n_kept * steps * len(rows) / t_learned, "unit": "ADO-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": 10, "ms_per_step": t_learned * 1e3 / (steps * len(rows)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded spin-glass parameter sets; network trained on {train_sets}, {len(rows)} held out)",
        "config": {"workload": "learned_closure_n16_L2", "spins": 4, "N": 16, "kept_tiers": "0..2",
                   "reference_tiers": "0..6", "network": "2 x 256 tanh, BF16 weights, FP32 accumulation (BF16 MFMA)"},
        "mean_err": {k: float(np.mean([r[f"err_{k}"] for r in rows])) for k in ("plain", "physics", "learned")},
        "time_to_solution_x_vs_deep_plain_fp64": t_deep / t_learned,
        "time_to_solution_x_vs_plain_fp64_at_matched_error": t_matched / t_learned,
        "held_out": rows,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_legacy(args):
    """The reference's compiled legacy spin glass (N = 1024, one bath with bit-flip coupling, K = 2,
    JMax = 3 -> 10 ADOs, dt 0.1, Taylor order 4), the only workload the reference publishes a time for."""
    import helix_amd as hx
    m = hx.Model.legacy(N=1024, precision=args.precision)
    plan = hx.Plan(m)
    plan.set_initial()
    plan.step(args.warmup)
    t0 = time.perf_counter()
    plan.step(args.steps)
    sec = time.perf_counter() - t0
    stage_ms, _ = plan.time_steps(max(1, args.timed_steps))
    value = m.H * args.steps / sec
    # the reference's published legacy time (RTX 4070 SUPER, FP32): 85.718 ms per step of its 10 ADOs
    ref_value = 10 / 85.718e-3
    line = {"metric": "ADO-steps/sec on 1 MI355X", "value": value, "unit": "ADO-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3 / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": value / ref_value if args.precision == "fp32" else None,
            "baseline_note": "reference FP32 legacy profile on an RTX 4070 SUPER: 85.718 ms/step = 116.7 ADO-steps/s",
            "dtype": "f64" if args.precision == "fp64" else "f32",
            "data": "legacy compiled model", "config": {"workload": "legacy_spin_glass_n1024", "n_ado": m.H,
                                                        "precision": args.precision},
            "stage_ms": [float(x) for x in stage_ms], "launches_per_step": plan.info()["launches_per_step"]}
    print(json.dumps(line), flush=True)
    plan.close()
    return 0


def bench_batched(args):
    """512 independent spin-glass instances (seeds 1000..1511: spins 3-6, tiers 2-5, random bath
    parameters, dt 0.01) advanced together, one launch per Taylor stage, step replayed from a hipGraph."""
    import helix_amd as hx
    models = [hx.Model.spin_glass_random(1000 + i, dt=0.01, precision=args.precision)
              for i in range(args.batch_count)]
    batch = hx.Batch(models, graph=True, autotune=True)
    batch.set_initial()
    info = batch.info()
    print(f"# batched: {len(models)} instances, {info['total_ado']} ADOs, {info['device_bytes'] / 2**30:.1f} GiB, "
          f"threads {info['threads']}", file=sys.stderr, flush=True)
    if args.warmup:
        batch.step(args.warmup)
    t0 = time.perf_counter()
    batch.step(args.steps)
    sec = time.perf_counter() - t0
    s_bytes = 16 if args.precision == "fp64" else 8
    herm = batch.layout() == "hermitian"
    alg_bytes_step = sum(14 * ((m.N * (m.N + 1)) // 2 if herm else m.N * m.N) * s_bytes * m.H for m in models)
    ms_step = sec * 1e3 / args.steps
    achieved = alg_bytes_step / (ms_step * 1e-3) / 1e9
    line = {
        "metric": "ADO-steps/sec on 1 MI355X", "value": info["total_ado"] * args.steps / sec, "unit": "ADO-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded random spin-glass instances)",
        "config": {"workload": f"batched_{len(models)}", "instances": len(models), "n_ado": info["total_ado"],
                   "precision": args.precision, "threads": info["threads"], "graph": info["graph"],
                   "launches_per_step": info["launches_per_step"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_batch_stage",
                     "basis": "14 P s bytes per ADO-step, P stored elements per ADO of the batch's "
                              + ("Hermitian" if herm else "full") + " layout"},
    }
    print(json.dumps(line), flush=True)
    batch.close()
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="spin_glass_n64_L5", choices=sorted(WORKLOADS) + ["batched", "learned", "legacy"])
    ap.add_argument("--batch-count", type=int, default=512, help="instances of the batched workload")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "split_fp32"])
    ap.add_argument("--timed-steps", type=int, default=100, help="steps timed per stage launch with HIP events")
    ap.add_argument("--dense-hamiltonian", action="store_true",
                    help="route -i[H, rho] through the dense MFMA commutator (N = 16, 32, 64)")
    ap.add_argument("--closure-only", action="store_true",
                    help="--workload learned: run only the learned solver (seed 9000) for --steps steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP32 / split-FP32 / batched lines measured beside the primary workload")
    ap.add_argument("--form-steps", type=int, default=500, help="timed steps of each single-problem form")
    ap.add_argument("--batch-steps", type=int, default=30, help="timed steps of each batched form")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the attainable-peak measurement (HBM read/copy, FP64/FP32 MFMA) reported in roofline")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (calibrates rocprof byte counters)")
    args = ap.parse_args(argv)

    if args.gpus != 1:  # every workload: the figure is one GPU's, and nothing here spans GPUs
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        return 2
    if args.workload.startswith("batched"):
        return bench_batched(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "legacy":
        return bench_legacy(args)

    import helix_amd as hx

    spec = WORKLOADS[args.workload]
    spins, seed, L, dt, lam, gamma, T, K = spec
    model = hx.Model.spin_glass(spins, seed, lam=lam, gamma=gamma, T=T, K=K, L=L, dt=dt, precision=args.precision)
    if args.dense_hamiltonian:
        model.use_dense_hamiltonian()
    plan = hx.Plan(model)
    plan.set_initial()
    print(f"# workload {args.workload}: N={model.N} ADOs={model.H} modes={model.M} precision={args.precision}",
          file=sys.stderr, flush=True)
    # per-stage launch times with HIP events on the plan's stream (the roofline's figure), measured before the timed
    # region: these untimed steps also bring the GPU to its steady clocks, which a short warmup alone does not (20
    # steps after 5 warmup steps ran 6% slower than after 200)
    stage_ms, total_ms = plan.time_steps(max(1, args.timed_steps))
    plan.set_initial()
    if args.warmup:
        plan.step(args.warmup)
    t0 = time.perf_counter()
    plan.step(args.steps)  # returns after the device finished
    t1 = time.perf_counter()
    sec = t1 - t0
    value = model.H * args.steps / sec

    s_bytes = 16 if args.precision == "fp64" else 8
    n2 = model.N * model.N
    info = plan.info()
    layout = plan.layout()
    stored = (model.N * (model.N + 1)) // 2 if layout == "hermitian" else n2
    alg_bytes_step = 14 * stored * s_bytes * model.H
    per_launch = alg_bytes_step / model.order
    full_per_launch = 14 * n2 * s_bytes * model.H / model.order
    mean_launch_ms = float(sum(stage_ms) / len(stage_ms))
    achieved = per_launch / (mean_launch_ms * 1e-3) / 1e9
    line = {
        "metric": "ADO-steps/sec on 1 MI355X",
        "value": value,
        "unit": "ADO-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded spin-glass couplings, rho_0 = |up..up><up..up|)",
        "config": {"workload": args.workload, "spins": spins, "N": model.N, "n_ado": model.H, "modes": model.M,
                   "max_tier": L, "dt": dt, "order": model.order, "precision": args.precision, "seed": seed,
                   "hamiltonian": "dense (MFMA)" if args.dense_hamiltonian else "spin-structured (diag + flips)",
                   "layout": layout},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "basis": f"14 P s bytes per ADO-step over the {layout} layout's P = "
                              f"{stored} stored elements per ADO, s = {s_bytes}",
                     "kernel": ("k_stage_dense" if args.dense_hamiltonian else "k_stage_tile")
                     + " (fused Liouville RHS + Taylor epilogue)",
                     "alg_bytes_per_launch": per_launch, "mean_launch_ms": mean_launch_ms,
                     "stage_ms": [float(x) for x in stage_ms],
                     "effective_full_block_GBs": full_per_launch / (mean_launch_ms * 1e-3) / 1e9},
        "launches_per_step": info["launches_per_step"],
    }
    if args.dense_hamiltonian:
        # the dense commutator is MFMA-bound: 8 N^3 flops per ADO and stage for a real H (H rho and rho H on the real
        # and imaginary planes of rho), 16 N^3 for a complex H (the 64 N^3 per ADO-step of the full complex GEMMs)
        hreal = bool(model.hamiltonian_is_real())
        flops_launch = (8 if hreal else 16) * model.N ** 3 * model.H * (2 if args.precision == "split_fp32" else 1)
        peak = 78.6 if args.precision == "fp64" else 157.3
        tflops = flops_launch / (mean_launch_ms * 1e-3) / 1e12
        line["roofline"] = {
            "bound": "mfma", "achieved": tflops, "peak": peak, "unit": "TFLOP/s", "frac": tflops / peak,
            "traffic": None,
            "basis": f"{8 if hreal else 16} N^3 flops per ADO and stage issued on MFMA (real H: the imaginary-H "
                     "products are skipped" + ("; split-FP32 issues them twice" if args.precision == "split_fp32" else "")
                     + f"); complex-GEMM count 16 N^3 gives {16 * model.N ** 3 * model.H / (mean_launch_ms * 1e-3) / 1e12:.1f}",
            "kernel": "k_stage_dense_tiled" if model.N > 64 else "k_stage_dense",
            "flops_per_launch": flops_launch, "mean_launch_ms": mean_launch_ms,
            "stage_ms": [float(x) for x in stage_ms],
            "hbm_GBs_stored_bytes": achieved}
    if not args.no_peaks:
        # what this GPU sustains, measured after the timed region by the library's own kernels: the fractions above
        # are against datasheet peaks; these put the same achieved figure against a copy kernel (HBM) or a dense
        # MFMA loop of the instruction the dense kernels issue
        pk = hx.measure_peaks(0, 1 << 30, 5)
        rf = line["roofline"]
        rf["attainable"] = pk
        if rf["bound"] == "hbm":
            rf["frac_of_measured_copy"] = rf["achieved"] / pk["hbm_copy_gbs"]
        else:
            rf["frac_of_measured_mfma"] = rf["achieved"] / pk["mfma_f64_tflops" if args.precision == "fp64"
                                                               else "mfma_f32_tflops"]
    traffic = None if args.dense_hamiltonian else _profiled_traffic(args.workload, args.precision, layout)
    if traffic is not None:
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_source"] = ("committed constant, not measured in this run: mean HBM bytes per stage launch "
                                              "from profiles/stage_traffic_spin_glass_n64_L5.json (rocprofv3 "
                                              "FETCH_SIZE/WRITE_SIZE of this kernel, scaled by a device copy)")
    if not args.no_forms and args.workload == "spin_glass_n64_L5" and args.precision == "fp64" \
            and not args.dense_hamiltonian:
        line["forms"] = measure_forms(args, value)
    if args.copy_probe_mib:
        line["copy_probe_bytes"] = copy_probe(args.copy_probe_mib << 20)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec)
    print(json.dumps(line), flush=True)
    plan.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Headline benchmark: site-updates/ns on one MI355X, batched FP32 vs plain FP64.

Workload (the XY benchmark configuration): XY model, J = 1, L = 128 periodic,
32 replicas on a geometric temperature ladder 0.7-1.1, angles uniform from
seed 170, Gaussian-MH angle proposals whose widths are adapted toward 0.5
acceptance per temperature (2,000 sweeps, then frozen), replica exchange
every 10 sweeps.  One step = 10 sweeps + 1 exchange.

* value: the batched FP32 engine (core.batched.BatchedRun, one persistent
  launch for all timed steps), site-updates per ns of wall-clock time with the
  state already resident on the GPU.  A site-update is one visit of one site
  in one colour pass (L*L*R per sweep).
* plain_fp64: the drop-in path (Runtime + ParallelTempering, float64 XY with
  the same Gaussian widths, four colour launches per sweep, queued by one
  native call per sweep); plain_fp64_python_launches: the same with the
  model's update_color called from Python for each colour, the reference's
  launch pattern.  speedup_vs_plain_fp64[_python_launches] = value / each.
* uniform_window: the batched engine with the reference's uniform window
  (delta = 1.5) on the same lattice, for comparison.
* roofline: algorithmic bytes of the batched kernel, 40 B per site-update
  (float32 angle read + write, 16 B xoroshiro128+ state read + write), over
  its HIP-event time, against 8 TB/s.  `traffic` is the HBM bytes of the timed
  launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE measurement
  (per sweep, scaled by a copy kernel of known size; profiles/) times the
  launch's sweeps, when that measurement used the same workload and strip height.
  The engine keeps lattice and streams on chip, so its real bound is VALU
  issue: `roofline.valu_issue` prices the timed launch against 256 CUs x 4
  SIMDs x 2.4 GHz / 4 cycles per wave64 instruction, at the VALU instructions
  per site-update counted by rocprofv3 for the same launch shape (profiles/).
* cpu_baseline: the CPU oracle's restatement of the same Gaussian update
  (float32, OpenMP over replicas) on the same lattice, a bounded number of
  sweeps on the host cores.
* fss_packed: the finite-size-scaling set in one packed launch
  (core.packed.PackedRun): Ising L in {16, 24, 32, 48, 64, 96, 128, 256} with
  48 temperatures 2.1-2.45 and XY (uniform window) L in {16, 32, 64, 128, 256}
  with 48 temperatures 0.8-1.0, exchanges every 10 sweeps, 8.96 M
  site-updates per sweep; its own roofline (34 B per Ising and 40 B per XY
  site-update).
* ising_config1: the Ising configuration (L = 32, 8 temperatures, exchange
  every sweep, 2,000 + 10,000 sweeps) as a whole run on the drop-in path, the
  batched engine and the CPU oracle (one thread), with their end states
  compared.
* gradients: the differentiable energy (XY and Kawamura, L = 256, 64
  replicas, float32 and float64), forward + backward per call, against 3 x
  the state bytes per call; measured in a child process, the only one that
  imports torch.
* films: the drop-in path of the Kawamura model (L = 192) and of the He3
  Gaussian sampler (128 x 128), 64 replicas each with on-device exchanges
  every 10 sweeps, in float32 and float64 (main/precision_models.py holds
  their ladders and parameters); algorithmic bytes 2 x state + 32 per
  site-update (Kawamura 72 / 112 B, He3 128 / 224 B).
* he3_reference_protocol: the reference's own published benchmark (He3
  Cartesian sampler, L = 32, 48, 64, 80, 100 replicas; updates only, with
  exchanges, and with a measurement every 10 sweeps) on the drop-in path,
  with the speedup over the reference's published ms per sweep.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build measures one GPU.
"""

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
for p in (str(REPO / "gpu-monte-carlo_amd"), str(REPO)):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "site-updates/ns on 1 GPU (batched FP32 vs plain FP64) + % HBM roofline"
L, R, SEED, EXCHANGE = 128, 32, 170, 10
ADAPT_SWEEPS, ADAPT_INTERVAL = 2000, 100
BYTES_PER_SITE_UPDATE = 40          # XY float32: 2 * 4 B angle + 2 * 16 B RNG state
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = REPO / "profiles" / "batch_kernel_xy_gauss_L128_R32.traffic.json"
VALU_FILE = REPO / "profiles" / "batch_kernel_xy_gauss_L128_R32.valu.json"
FSS_ISING_L = (16, 24, 32, 48, 64, 96, 128, 256)
FSS_XY_L = (16, 32, 64, 128, 256)
FSS_R = 48
FSS_CONFIG = (512, 2048)            # workgroup size, task sites (autotuned on MI355X)


def temperatures():
    return np.geomspace(0.7, 1.1, R)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _model(proposal, sigma=1.0, dtype=np.float32):
    from models.xy.model import XYModel
    if proposal == "gaussian":
        return XYModel(J=1.0, dtype=dtype, start="random", proposal="gaussian", sigma=sigma)
    return XYModel(J=1.0, delta=1.5, dtype=dtype, start="random")


def bench_batched(steps, warmup, strip_rows, tune, proposal="gaussian", threads=None):
    from core.batched import BatchedRun, autotune
    from mcgpu import synchronize
    from mcgpu.timing import EventTimer

    model = _model(proposal)
    if strip_rows is None and tune:
        log(f"autotuning strip height and workgroup size ({proposal})")
        strip_rows, threads = autotune(model, L, temperatures(), seed=SEED, exchange_interval=EXCHANGE,
                                       sweeps=400, verbose=True)
    run = BatchedRun(model, L, temperatures(), seed=SEED, exchange_interval=EXCHANGE, strip_rows=strip_rows,
                     threads=threads)
    sigmas = None
    if proposal == "gaussian":
        log(f"adapting Gaussian widths over {ADAPT_SWEEPS} sweeps")
        run.adapt_sigmas(ADAPT_SWEEPS, interval=ADAPT_INTERVAL)
        sigmas = run._sigmas_host.copy()
    log(f"batched FP32 ({proposal}): strip rows {run.H}, {run.R * run.strips} workgroups of {run.threads} "
        f"threads; warmup {warmup} steps")
    run.run(warmup * EXCHANGE)
    run.reset_acceptance()
    synchronize()
    timer = EventTimer(run._stream)
    sweeps = steps * EXCHANGE
    log(f"timing {steps} steps ({sweeps} sweeps) in one launch")
    t0 = time.perf_counter()
    timer.start()
    handle = run.launch(sweeps)
    timer.stop()
    run.collect(handle)
    synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = timer.elapsed_ms()
    energy = run.measure()["energy"]
    if not np.all(np.isfinite(energy)):
        raise RuntimeError("batched run produced nonfinite energies")
    d = run.diagnostics()
    return {
        "wall_s": wall,
        "kernel_ms": kernel_ms,
        "site_updates": sweeps * L * L * R,
        "strip_rows": run.H,
        "threads": run.threads,
        "workgroups": run.R * run.strips,
        "pt_acceptance": float(d["acceptance_rates"].mean()),
        "move_acceptance": d["acceptance_by_temperature"],
        "sigmas": sigmas,
    }


def bench_plain(dtype, sweeps, sigmas, warmup=20, python_launches=False):
    """The drop-in path: Runtime + ParallelTempering, one pt.sweep() call per sweep (four colour
    launches).  By default a sweep is one native call that queues the four launches; with
    python_launches the model's update_color is called from Python for each colour, as the
    reference's Runtime.sweep does (the per-colour plain path, core/runtime.py:87-99)."""
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from mcgpu import synchronize

    model = _model("gaussian", sigmas, dtype)
    if python_launches:
        model.update_color = model.update_color  # a method set on the instance: no native chain
    rt = Runtime(model, L, temperatures(), seed=SEED)
    pt = ParallelTempering(rt, exchange_interval=EXCHANGE, seed=SEED + 1)
    for _ in range(warmup):
        pt.sweep()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(sweeps):
        pt.sweep()
    synchronize()
    wall = time.perf_counter() - t0
    return {"wall_s": wall, "site_updates": sweeps * L * L * R, "sweeps": sweeps}


def fss_ensembles():
    from core.batched import Ensemble
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    ising_T = np.linspace(2.1, 2.45, FSS_R)
    xy_T = np.linspace(0.8, 1.0, FSS_R)
    ens = [Ensemble(IsingModel(start="random"), n, ising_T, seed=42 + n, exchange_interval=EXCHANGE, pt_seed=43 + n)
           for n in FSS_ISING_L]
    ens += [Ensemble(XYModel(start="random"), n, xy_T, seed=42 + n, exchange_interval=EXCHANGE, pt_seed=43 + n)
            for n in FSS_XY_L]
    return ens


def bench_fss(sweeps, warmup=50, config=FSS_CONFIG):
    """The finite-size-scaling set in one packed launch (HIP-event and wall time)."""
    from core.packed import PackedRun
    from mcgpu.timing import EventTimer
    ens = fss_ensembles()
    pr = PackedRun(ens, threads=config[0], task_sites=config[1])
    sites = sum(e.L * e.L * e.R for e in ens)
    ising_sites = sum(e.L * e.L * e.R for e in ens if e.code == 0)
    log(f"packed finite-size set: {len(ens)} ensembles, {sites} sites, {pr.workgroups} workgroups of "
        f"{pr.threads} threads, {pr.plan['tasks_per_workgroup']} tasks each")
    pr.run(warmup)
    pr.synchronize()
    timer = EventTimer(pr._stream)
    t0 = time.perf_counter()
    timer.start()
    handle = pr.launch(sweeps)
    timer.stop()
    pr.collect(handle)
    wall = time.perf_counter() - t0
    kernel_ms = timer.elapsed_ms()
    algo = sweeps * (34 * ising_sites + 40 * (sites - ising_sites))
    achieved = algo / (kernel_ms * 1e-3) / 1e9
    rate = sites * sweeps / (wall * 1e9)
    log(f"packed finite-size set: {rate:.2f} site-updates/ns wall, {kernel_ms / sweeps * 1e3:.2f} us/sweep in-kernel")
    return {
        "value": round(rate, 4), "unit": "site-updates/ns",
        "kernel_site_updates_per_ns": round(sites * sweeps / (kernel_ms * 1e6), 3),
        "us_per_sweep": round(kernel_ms * 1e3 / sweeps, 3),
        "sweeps": sweeps, "site_updates_per_sweep": sites,
        "workgroups": pr.workgroups, "workgroup_threads": pr.threads, "task_sites": pr.task_sites,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": algo},
        "workload": (f"Ising L={list(FSS_ISING_L)} x {FSS_R} T (2.1-2.45) and XY L={list(FSS_XY_L)} x {FSS_R} "
                     f"T (0.8-1.0), random starts, exchange every {EXCHANGE} sweeps, one packed launch"),
    }


def bench_film(kind, precise, sweeps, warmup=20):
    """Drop-in path throughput of a film model with on-device PT every 10 sweeps."""
    from core.parallel_tempering import ParallelTempering
    from main.precision_models import runtime, temperatures
    from mcgpu import synchronize
    from mcgpu.timing import EventTimer
    R = 64
    L = 192 if kind == "kawamura" else 128
    rt = runtime(kind, L, temperatures(kind, R), precise, 42 + L)
    pt = ParallelTempering(rt, exchange_interval=EXCHANGE, seed=43 + L)
    for _ in range(warmup):
        pt.sweep()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(sweeps):
        pt.sweep()
    synchronize()
    wall = time.perf_counter() - t0
    state_bytes = {("kawamura", False): 20, ("kawamura", True): 40, ("he3", False): 48, ("he3", True): 96}
    b = 2 * state_bytes[(kind, precise)] + 32
    rate = L * L * R * sweeps / (wall * 1e9)
    return {"value": round(rate, 4), "unit": "site-updates/ns", "L": L, "replicas": R, "sweeps": sweeps,
            "dtype": ("f64" if precise else "f32") if kind == "kawamura" else ("c128" if precise else "c64"),
            "ms_per_sweep": round(wall * 1e3 / sweeps, 4), "bytes_per_site_update": b,
            "roofline_frac": round(rate * b / HBM_PEAK_GBS, 4)}


# The reference's published He3 benchmark (profiling/HE3_BENCHMARK_20261027.md:9-38, RTX 4070 SUPER,
# 100 replicas, medians of 5 windows x 3,000 sweeps after 1,000 warm-up sweeps), ms per sweep by L
HE3_REF_MS = {
    "updates": {32: 0.166939, 48: 0.140702, 64: 0.156559, 80: 0.256459},
    "pt": {32: 0.166103, 48: 0.160599, 64: 0.160094, 80: 0.257237},
    "production": {32: 0.199653, 48: 0.183814, 64: 0.240163, 80: 0.317248},
}


def bench_he3_reference(sizes=(32, 48, 64, 80), replicas=100, warmup=1000, sweeps=3000, repeats=5):
    """The reference's He3 benchmark protocol (profiling/profile_he3.py:49-140) on the drop-in path:
    He3Model(K=1, delta=0.15, pressure=20, reduced_temperature=1), the ladder
    make_temperature_ladder(0.15, 0.45, 100, Tc=0.257, tc_window=0.06, tc_fraction=0.6),
    Runtime seed 42 + L, ParallelTempering every 10 sweeps with seed 43 + L; modes "updates"
    (Runtime.sweep), "pt" (ParallelTempering.sweep) and "production" (pt plus a measurement copied
    to the host every 10 sweeps); the median of `repeats` windows of `sweeps` sweeps, against the
    reference's published ms per sweep."""
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from mcgpu import synchronize
    from models.he3.model import He3Model
    from utils.temperature_ladder import make_temperature_ladder
    temps = make_temperature_ladder(T_min=0.15, T_max=0.45, n_T=replicas, Tc=0.257, tc_window=0.06,
                                    tc_fraction=0.60)
    out = {"workload": (f"He3 Cartesian Metropolis (K=1, delta=0.15, P=20 bar, t=1), {replicas} replicas, "
                        f"exchange every 10 sweeps, median of {repeats} windows of {sweeps} sweeps after "
                        f"{warmup} warm-up sweeps (the reference's published protocol)"),
           "reference_hardware": "RTX 4070 SUPER (profiling/HE3_BENCHMARK_20261027.md)", "cases": []}
    for Ln in sizes:
        rt = Runtime(He3Model(K=1, delta=0.15, pressure=20, reduced_temperature=1), Ln, temps, seed=42 + Ln)
        pt = ParallelTempering(rt, exchange_interval=10, seed=43 + Ln)
        for _ in range(warmup):
            pt.sweep()
        pt.measure()
        synchronize()
        for mode in ("updates", "pt", "production"):
            times = []
            for _ in range(repeats):
                synchronize()
                t0 = time.perf_counter()
                for step in range(1, sweeps + 1):
                    if mode == "updates":
                        rt.sweep()
                    else:
                        pt.sweep()
                        if mode == "production" and step % 10 == 0:
                            for v in pt.measure().values():
                                np.asarray(v).copy()
                synchronize()
                times.append(time.perf_counter() - t0)
            ms = float(np.median(times)) * 1e3 / sweeps
            ref = HE3_REF_MS[mode][Ln]
            out["cases"].append({"L": Ln, "mode": mode, "ms_per_sweep": round(ms, 5),
                                 "site_updates_per_ns": round(Ln * Ln * replicas / (ms * 1e6), 4),
                                 "reference_ms_per_sweep": ref, "speedup_vs_reference": round(ref / ms, 2)})
            log(f"he3 reference protocol L={Ln} {mode}: {ms:.4f} ms/sweep ({ref / ms:.2f}x the reference)")
        del pt, rt
    return out


def bench_ising_config1(therm=2000, meas=10000, interval=10):
    """The Ising configuration (L = 32, 8 temperatures 2.0-2.6, random start, exchange every sweep,
    2,000 + 10,000 sweeps, a sample every 10) as a whole run on three paths: the drop-in
    Simulation over ParallelTempering, the batched engine (one launch), and the CPU oracle
    (single thread, the same sweeps and exchanges).  The three runs end in identical spins."""
    import oracle
    from core.batched import BatchedRun, BatchedSimulation
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from core.simulation import Simulation
    from mcgpu import synchronize
    from models.ising.model import IsingModel
    Li, Ri = 32, 8
    temps = np.linspace(2.0, 2.6, Ri)
    seed, pt_seed = 42 + Li, 43
    sweeps = therm + meas
    out = {"workload": (f"Ising J=1 L={Li}, {Ri} temperatures 2.0-2.6, random start seed {seed}, exchange every "
                        f"sweep (PT seed {pt_seed}), {therm} + {meas} sweeps, a sample every {interval}"),
           "unit": "site-updates/ns", "site_updates": Li * Li * Ri * sweeps}
    model = IsingModel(start="random")
    rt = Runtime(model, Li, temps, seed=seed)
    pt = ParallelTempering(rt, exchange_interval=1, seed=pt_seed)
    synchronize()
    t0 = time.perf_counter()
    Simulation(pt, thermalization_sweeps=therm, measurement_sweeps=meas, sample_interval=interval).run()
    synchronize()
    wall = time.perf_counter() - t0
    out["plain"] = {"value": round(out["site_updates"] / (wall * 1e9), 4), "wall_s": round(wall, 3),
                    "path": "Simulation + ParallelTempering + Runtime, 4 colour launches per sweep"}
    br = BatchedRun(IsingModel(start="random"), Li, temps, seed=seed, exchange_interval=1, pt_seed=pt_seed,
                    strip_rows=8)
    synchronize()
    t0 = time.perf_counter()
    BatchedSimulation([br], thermalization_sweeps=therm, measurement_sweeps=meas, sample_interval=interval).run()
    wall = time.perf_counter() - t0
    out["batched"] = {"value": round(out["site_updates"] / (wall * 1e9), 4), "wall_s": round(wall, 3),
                      "path": f"BatchedSimulation, one launch of {Li // 8 * Ri} workgroups (strips of 8 rows)"}
    same = np.array_equal(br.snapshot(), rt.snapshot())
    oracle.set_threads(1)
    spins = model.initialize_state(Li, Ri, seed).copy()
    rng = oracle.rng_init(Ri * Li * Li, seed)
    betas = (1.0 / temps).astype(np.float32)
    w2t = np.arange(Ri, dtype=np.int32)
    t2w = np.arange(Ri, dtype=np.int32)
    counts = np.zeros((3, Ri - 1), np.int64)
    trips = np.zeros((5, Ri), np.int64)
    trips[0] = trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, Ri - 1
    trips[2, [0, -1]] = 0
    pt_rng = oracle.rng_init(Ri - 1, pt_seed)
    parity = 0
    t0 = time.perf_counter()
    for sweep in range(1, sweeps + 1):
        oracle.ising_sweeps(spins, rng, betas, model.J, 1)
        e = oracle.ising_energy(spins, model.J)
        oracle.pt_exchange(e, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep)
        parity = 1 - parity
        if sweep > therm and (sweep - therm) % interval == 0:
            oracle.ising_measure(spins, model.J)
    wall = time.perf_counter() - t0
    out["cpu_oracle"] = {"value": round(out["site_updates"] / (wall * 1e9), 5), "wall_s": round(wall, 3),
                         "cores": 1, "kind": "port"}
    out["end_states_identical"] = bool(same and np.array_equal(spins, rt.snapshot()))
    log(f"Ising config 1: plain {out['plain']['value']}, batched {out['batched']['value']}, "
        f"CPU {out['cpu_oracle']['value']} site-updates/ns; identical end states: {out['end_states_identical']}")
    return out


def gradients_in_child():
    """The differentiable-energy line in a child process: it runs through torch, whose HIP
    runtime must then be the process's only one, and a launch through torch's runtime costs this
    process's engine lines about 20 us more per launch than ROCm's own runtime."""
    import subprocess
    res = subprocess.run([sys.executable, "-u", str(Path(__file__).resolve()), "--gradients-only"],
                         capture_output=True, text=True)
    sys.stderr.write(res.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"the gradients line failed (exit {res.returncode})")
    return json.loads(res.stdout.strip().splitlines()[-1])


def bench_gradients(calls=20):
    """Differentiable energy at L = 256, 64 replicas: H(theta; J, twist) forward and dH/dstate,
    dH/dtwist, dH/dJ backward through the HIP kernels, XY and Kawamura, float32 and float64.
    ms_per_call is the whole autograd call (torch bookkeeping included); kernel_ms is the terms
    and gradient kernels queued back to back.  Algorithmic bytes per call: the state read by the
    terms kernel, read again and the gradient written by the gradient kernel (3 x state bytes)."""
    import torch
    import autograd
    from autograd import lattice_energy
    from models.kawamura.model import KawamuraModel
    from models.xy.model import XYModel
    Lg, Rg = 256, 64
    out = {}
    for name in ("xy", "kawamura"):
        for dt in (torch.float32, torch.float64):
            npdt = np.float32 if dt == torch.float32 else np.float64
            m = XYModel(start="random", dtype=npdt) if name == "xy" else KawamuraModel(dtype=npdt, start="random")
            host = np.ascontiguousarray(m.initialize_state(Lg, Rg, 7))
            state = torch.tensor(host, device="cuda")
            J = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
            twist = torch.zeros(Rg, dtype=torch.float64, device="cuda", requires_grad=True)

            def once():
                x = state.detach().requires_grad_(True)
                lattice_energy(x, J, twist, name).sum().backward()
                return x.grad

            for _ in range(3):
                once()
            torch.cuda.synchronize()
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(calls):
                once()
            stop.record()
            torch.cuda.synchronize()
            ms = start.elapsed_time(stop) / calls
            # the kernels alone: terms (forward) and state gradient (backward) through the C-ABI on
            # preallocated buffers, queued back to back
            from mcgpu import _lib
            terms_fn, grad_fn = autograd._KERNELS[name][0], autograd._KERNELS[name][1]
            code, stream, ptr = autograd._code(state), autograd._stream(), autograd._ptr
            tw = torch.zeros(Rg, dtype=torch.float64, device="cuda")
            ones = torch.ones(Rg, dtype=torch.float64, device="cuda")
            zeros = torch.zeros(Rg, dtype=torch.float64, device="cuda")
            out3 = torch.empty((3, Rg), dtype=torch.float64, device="cuda")
            gout = torch.empty_like(state)

            def kernels():
                _lib.call(terms_fn, stream, code, ptr(state), Rg, Lg, ptr(tw), ptr(out3))
                _lib.call(grad_fn, stream, code, ptr(state), Rg, Lg, ptr(tw), ptr(ones), ptr(ones), ptr(zeros),
                          ptr(gout))

            kernels()
            torch.cuda.synchronize()
            start.record()
            for _ in range(calls):
                kernels()
            stop.record()
            torch.cuda.synchronize()
            kms = start.elapsed_time(stop) / calls
            nbytes = 3 * state.numel() * state.element_size()
            key = f"{name}_{'f32' if dt == torch.float32 else 'f64'}"
            out[key] = {"ms_per_call": round(ms, 4), "kernel_ms": round(kms, 4),
                        "GB_per_s": round(nbytes / (kms * 1e-3) / 1e9, 1),
                        "roofline_frac": round(nbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes": nbytes}
            log(f"gradients {key}: {ms:.3f} ms per autograd forward+backward, kernels {kms:.4f} ms "
                f"({out[key]['GB_per_s']} GB/s)")
    out["workload"] = (f"lattice_energy(state, J, twist) forward + backward (dH/dstate, dH/dtwist, dH/dJ), L={Lg}, "
                       f"{Rg} replicas, random angles")
    return out


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB: reads and writes exactly mib MiB."""
    from mcgpu import _lib, as_pointer, device_zeros, synchronize
    n = mib * (1 << 20)
    src = device_zeros(n, np.uint8)
    dst = device_zeros(n, np.uint8)
    synchronize()
    _lib.call("mc_memcpy_dtod", as_pointer(dst), as_pointer(src), n)
    synchronize()
    log(f"copy probe: {mib} MiB device-to-device")


def measured_traffic(strip_rows, threads, sweeps):
    """HBM bytes of a launch of `sweeps` sweeps from the committed counter runs of the same
    strip height and workgroup size, or None."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None, None
    for c in t.get("configs", []):
        if int(c["strip_rows"]) == strip_rows and int(c["threads"]) == threads:
            return float(c["traffic_bytes_per_sweep"]) * sweeps, c.get("summary")
    return None, None


def valu_roofline(strip_rows, threads, kernel_rate):
    """The VALU-issue bound of the batched kernel from the committed instruction counters of the
    same launch shape: wave-instructions per second of the chip x 64 lanes / VALU instructions per
    site-update, and the fraction of it the timed launch reached (None without matching counters)."""
    try:
        v = json.loads(VALU_FILE.read_text())
    except (OSError, ValueError):
        return None
    if int(v["strip_rows"]) != strip_rows or int(v["threads"]) != threads:
        return None
    peak = float(v["peak_wave_instructions_per_s"]) * 64 / float(v["valu_instructions_per_site_update"]) / 1e9
    return {"unit": "site-updates/ns", "peak": round(peak, 1), "achieved": round(kernel_rate, 2),
            "frac": round(kernel_rate / peak, 4),
            "valu_instructions_per_site_update": v["valu_instructions_per_site_update"],
            "source": v["summary"]}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(target_s, sigmas):
    """The oracle's Gaussian XY update on the host cores (OpenMP over replicas, at most 16
    threads), then on one thread for a third of the time."""
    import oracle
    theta = _model("gaussian").initialize_state(L, R, SEED)
    rng = oracle.rng_init(R * L * L, SEED)
    betas = (1.0 / temperatures()).astype(np.float32)
    sig = np.asarray(sigmas, dtype=np.float32)

    def sweep():
        for color in range(4):
            oracle.xy_gauss_color(theta, rng, betas, color, 1.0, betas, sig)

    def measure(threads, seconds):
        oracle.set_threads(threads)
        sweep()  # warm caches and threads
        sweeps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            for _ in range(EXCHANGE):
                sweep()
            oracle.xy_energy(theta, 1.0)
            sweeps += EXCHANGE
        wall = time.perf_counter() - t0
        return sweeps * L * L * R / (wall * 1e9), sweeps, wall

    threads = min(16, os.cpu_count() or 1)
    value, sweeps, wall = measure(threads, target_s)
    one, sweeps1, wall1 = measure(1, target_s / 3)
    return {
        "value": value,
        "unit": "site-updates/ns",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sweeps} sweeps of the Gaussian-proposal XY float32 update (C restatement of the "
                   f"engine's plain path, OpenMP over replicas) on L={L}, R={R} with the adapted widths, "
                   f"plus a float64 energy pass every {EXCHANGE} sweeps; {wall:.1f} s"),
        "single_thread": {"value": one, "sample": f"{sweeps1} sweeps on one thread; {wall1:.1f} s"},
        "host": {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count()},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000, help="timed steps (10 sweeps + 1 exchange each)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--plain-sweeps", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-uniform", action="store_true")
    ap.add_argument("--strip-rows", type=int, default=None)
    ap.add_argument("--threads", type=int, default=None, help="workgroup size with --strip-rows (256 or 512)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--fss-sweeps", type=int, default=1000, help="timed sweeps of the packed finite-size set")
    ap.add_argument("--no-fss", action="store_true")
    ap.add_argument("--film-sweeps", type=int, default=200, help="timed sweeps of each film-model path")
    ap.add_argument("--no-films", action="store_true")
    ap.add_argument("--no-ising", action="store_true", help="skip the Ising L=32 R=8 whole-run comparison")
    ap.add_argument("--no-gradients", action="store_true", help="skip the differentiable-energy line")
    ap.add_argument("--no-he3-ref", action="store_true", help="skip the reference's He3 benchmark protocol")
    ap.add_argument("--he3-sizes", default="32,48,64,80", help="lattice sizes of the He3 protocol (comma list)")
    ap.add_argument("--he3-warmup", type=int, default=1000, help="warm-up sweeps of the He3 protocol")
    ap.add_argument("--he3-sweeps", type=int, default=3000, help="sweeps per timed window of the He3 protocol")
    ap.add_argument("--he3-repeats", type=int, default=5, help="timed windows per mode of the He3 protocol")
    ap.add_argument("--gradients-only", action="store_true",
                    help="print the differentiable-energy line alone (bench.py runs it in a child process)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device once, so a counter run holds a "
                         "plain copy kernel of known size to scale its byte counters by")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the engine shards by "
              "replica/lattice size across GPUs by design but this build runs and reports N=1.",
              file=sys.stderr)
        sys.exit(2)

    if args.gradients_only:
        # torch's HIP runtime must be the process's only one: load it before
        # libmcgpu binds to a runtime (the gradients line runs through torch)
        import torch  # noqa: F401
        if args.copy_probe_mib > 0:
            copy_probe(args.copy_probe_mib)
        print(json.dumps(bench_gradients()), flush=True)
        return
    # Python's full garbage collections traverse every live object; with
    # torch loaded one takes ~40 ms on the box's host, and the drop-in path's
    # per-sweep calls trigger them now and then inside a timed loop.  The
    # objects that exist after setup are moved out of the collector's scope.
    gc.collect()
    gc.freeze()
    from mcgpu.timing import device_info
    name, cus = device_info()
    log(f"device {name}, {cus} CUs")
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    b = bench_batched(args.steps, args.warmup, args.strip_rows, not args.no_autotune, threads=args.threads)
    value = b["site_updates"] / (b["wall_s"] * 1e9)
    kernel_rate = b["site_updates"] / (b["kernel_ms"] * 1e6)
    achieved = BYTES_PER_SITE_UPDATE * b["site_updates"] / (b["kernel_ms"] * 1e-3) / 1e9
    log(f"batched FP32: {value:.2f} site-updates/ns wall, {kernel_rate:.2f} in-kernel")
    traffic, traffic_src = measured_traffic(b["strip_rows"], b["threads"], args.steps * EXCHANGE)

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "site-updates/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b["wall_s"] * 1e3 / args.steps, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"XY L={L} R={R}, T geometric 0.7-1.1, random start seed {SEED}, Gaussian-MH "
                         f"proposals with widths adapted toward 0.5 acceptance ({ADAPT_SWEEPS} sweeps, then "
                         f"frozen), exchange every {EXCHANGE} sweeps; batched FP32 persistent launch"),
            "L": L, "replicas": R, "sweeps_per_step": EXCHANGE,
            "strip_rows": b["strip_rows"], "workgroup_threads": b["threads"], "workgroups": b["workgroups"],
            "device": name, "compute_units": cus,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic),
            "traffic_source": traffic_src,
            "algorithmic_bytes": BYTES_PER_SITE_UPDATE * b["site_updates"],
            "bytes_per_site_update": BYTES_PER_SITE_UPDATE,
            "kernel_ms": round(b["kernel_ms"], 3),
            "kernel_site_updates_per_ns": round(kernel_rate, 3),
            "valu_issue": valu_roofline(b["strip_rows"], b["threads"], kernel_rate),
        },
        "move_acceptance": [round(float(x), 4) for x in b["move_acceptance"]],
        "sigmas": [round(float(x), 4) for x in b["sigmas"]],
        "acceptance_rate_pt": round(b["pt_acceptance"], 4),
    }
    if not args.no_uniform:
        u = bench_batched(max(args.steps // 3, 1), max(args.warmup // 3, 1), args.strip_rows, not args.no_autotune,
                          proposal="uniform", threads=args.threads)
        out["uniform_window"] = {
            "value": round(u["site_updates"] / (u["wall_s"] * 1e9), 4), "unit": "site-updates/ns",
            "kernel_site_updates_per_ns": round(u["site_updates"] / (u["kernel_ms"] * 1e6), 3),
            "strip_rows": u["strip_rows"], "workgroup_threads": u["threads"],
            "note": "batched FP32 with the reference's uniform window, delta = 1.5"}
    if not args.no_plain:
        p64 = bench_plain(np.float64, args.plain_sweeps, b["sigmas"])
        v64 = p64["site_updates"] / (p64["wall_s"] * 1e9)
        p32 = bench_plain(np.float32, args.plain_sweeps, b["sigmas"])
        v32 = p32["site_updates"] / (p32["wall_s"] * 1e9)
        pp = bench_plain(np.float64, args.plain_sweeps, b["sigmas"], python_launches=True)
        vpp = pp["site_updates"] / (pp["wall_s"] * 1e9)
        log(f"plain FP64 {v64:.3f} (colour launches from Python {vpp:.3f}), plain FP32 {v32:.3f} site-updates/ns")
        out["plain_fp64"] = {"value": round(v64, 4), "unit": "site-updates/ns",
                             "ms_per_sweep": round(p64["wall_s"] * 1e3 / p64["sweeps"], 4),
                             "path": "Runtime + ParallelTempering, one native call per sweep queueing its "
                                     "4 colour launches"}
        out["plain_fp64_python_launches"] = {
            "value": round(vpp, 4), "unit": "site-updates/ns",
            "ms_per_sweep": round(pp["wall_s"] * 1e3 / pp["sweeps"], 4),
            "path": "Runtime + ParallelTempering, the model's update_color called from Python for each of "
                    "the 4 colours (the reference's launch pattern)"}
        out["plain_fp32"] = {"value": round(v32, 4), "unit": "site-updates/ns",
                             "ms_per_sweep": round(p32["wall_s"] * 1e3 / p32["sweeps"], 4)}
        out["speedup_vs_plain_fp64"] = round(value / v64, 2)
        out["speedup_vs_plain_fp64_python_launches"] = round(value / vpp, 2)
    if not args.no_fss:
        out["fss_packed"] = bench_fss(args.fss_sweeps)
    if not args.no_films:
        films = {}
        for kind in ("kawamura", "he3"):
            for precise in (False, True):
                r = bench_film(kind, precise, args.film_sweeps)
                log(f"{kind} {r['dtype']}: {r['value']:.2f} site-updates/ns, {r['ms_per_sweep']} ms per sweep")
                films[f"{kind}_{r['dtype']}"] = r
        out["films"] = films
    if not args.no_he3_ref:
        out["he3_reference_protocol"] = bench_he3_reference(
            sizes=tuple(int(x) for x in args.he3_sizes.split(",")), warmup=args.he3_warmup,
            sweeps=args.he3_sweeps, repeats=args.he3_repeats)
    if not args.no_ising:
        out["ising_config1"] = bench_ising_config1()
    if not args.no_gradients:
        out["gradients"] = gradients_in_child()
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, b["sigmas"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Float32 against float64 sweeps of the Kawamura and He3 models at film sizes.

* Kawamura: L = 192, random start (random theta, tau = +-1, unit d), 64
  temperatures uniform in beta over [0.65, 0.80] J -- the reference
  campaign's range (main/kawamura.py:46-55) -- uniform-window Metropolis
  (delta = 1.5).
* He3: 128 x 128, He3Model(pressure=20, g_d=0.001) with the reference
  campaign's coefficients (main/he3.py:23-36), the Gaussian sampler with
  rho = -0.9, 64 temperatures of the reference's ladder focused at
  T = 0.385 (main/he3.py:49-56).

Each model runs two independent chains through the drop-in path (Runtime or
He3GaussianRuntime + ParallelTempering with on-device exchanges every 10
sweeps + Simulation): float32 (complex64 for He3) and float64 (complex128),
with different seeds.  Every estimator of utils.analysis (kawamura_analysis:
energy, chirality, phase and spin order parameters with their Binder
cumulants, helicity; he3_analysis: energy, density, chirality, phase order,
their Binder cumulants, helicity) gets a blocked-jackknife value and error
per temperature, with blocks of at least 4 tau_int of the slowest series, and
the z-score of the float32 - float64 difference.  Each chain's throughput
(site-updates per ns of wall time, sampling and exchanges included) is
recorded.

    python -m main.precision_models --model both --out data/precision_models.json
"""

import argparse
import json
import time
from pathlib import Path

import numpy as np

from main.precision_study import block_size_for, tau_int

KAWAMURA = {"L": 192, "T": (0.65, 0.80), "delta": 1.5}
HE3 = {"L": 128, "pressure": 20.0, "g_d": 0.001, "rho": -0.9,
       "ladder": {"T_min": 0.365, "T_max": 0.415, "Tc": 0.385, "tc_window": 0.01, "tc_fraction": 0.35}}


def temperatures(kind, R):
    from utils.temperature_ladder import make_temperature_ladder
    if kind == "kawamura":
        return make_temperature_ladder(KAWAMURA["T"][0], KAWAMURA["T"][1], R)
    return make_temperature_ladder(n_T=R, **HE3["ladder"])


def runtime(kind, L, temps, precise, seed):
    """The drop-in runtime of one chain (float64 when `precise`)."""
    from core.runtime import Runtime
    if kind == "kawamura":
        from models.kawamura.model import KawamuraModel
        model = KawamuraModel(J=1.0, delta=KAWAMURA["delta"], dtype=np.float64 if precise else np.float32,
                              start="random")
        return Runtime(model, L, temps, seed=seed)
    from models.he3.model import He3Model
    from models.he3.sampling import He3GaussianRuntime
    model = He3Model(pressure=HE3["pressure"], g_d=HE3["g_d"], dtype=np.complex128 if precise else np.complex64)
    return He3GaussianRuntime(model, L, temps, seed=seed, rho=HE3["rho"])


def chain(kind, L, R, precise, thermalization, measurement, sample_interval, seed):
    from core.parallel_tempering import ParallelTempering
    from core.simulation import Simulation
    from mcgpu import synchronize
    temps = temperatures(kind, R)
    rt = runtime(kind, L, temps, precise, seed)
    pt = ParallelTempering(rt, exchange_interval=10, seed=seed + 1)
    t0 = time.perf_counter()
    run = Simulation(pt, thermalization_sweeps=thermalization, measurement_sweeps=measurement,
                     sample_interval=sample_interval).run()
    synchronize()
    wall = time.perf_counter() - t0
    sweeps = thermalization + measurement
    run["metadata"]["rate_site_updates_per_ns"] = L * L * R * sweeps / (wall * 1e9)
    return run


def _slow_series(kind, run, L):
    s = run["samples"]
    if kind == "kawamura":
        return {"energy": s["energy"], "chirality2": np.asarray(s["chirality"], float)**2,
                "spin2": s["spin2"]}
    chir = np.asarray(s["chirality"], float) / np.asarray(s["density"], float)
    return {"energy": s["energy"], "chirality": np.abs(chir),
            "phase2": np.asarray(s["phase_real"], float)**2 + np.asarray(s["phase_imag"], float)**2}


def taus(kind, run, L):
    """Largest tau_int over temperatures of each slow series, in samples."""
    return {name: max(tau_int(np.asarray(x, float)[:, t]) for t in range(np.asarray(x).shape[1]))
            for name, x in _slow_series(kind, run, L).items()}


def compare(kind, single, double, L):
    from utils.analysis import he3_analysis, kawamura_analysis
    analysis = kawamura_analysis if kind == "kawamura" else he3_analysis
    ts, td = taus(kind, single, L), taus(kind, double, L)
    n = min(len(single["samples"]["energy"]), len(double["samples"]["energy"]))
    block = block_size_for({**{f"f32_{k}": v for k, v in ts.items()}, **{f"f64_{k}": v for k, v in td.items()}}, n)
    a32, a64 = analysis(single, L, block), analysis(double, L, block)
    out = {"block_size_samples": block, "tau_int_samples": {"float32": ts, "float64": td}, "estimators": {}}
    for name in (k for k in a32 if not k.endswith("_err")):
        v32, e32 = np.asarray(a32[name], float), np.asarray(a32[name + "_err"], float)
        v64, e64 = np.asarray(a64[name], float), np.asarray(a64[name + "_err"], float)
        z = (v32 - v64) / np.sqrt(e32**2 + e64**2)
        finite = np.isfinite(z)
        out["estimators"][name] = {
            "float32": v32.tolist(), "float32_err": e32.tolist(),
            "float64": v64.tolist(), "float64_err": e64.tolist(), "z": z.tolist(),
            "mean_z2": float(np.mean(z[finite]**2)) if finite.any() else None,
            "max_abs_z": float(np.max(np.abs(z[finite]))) if finite.any() else None,
        }
    return out


def run(kinds=("kawamura", "he3"), R=64, thermalization=10000, measurement=20000, sample_interval=10,
        seed=42, sizes=None):
    report = {"replicas": R, "thermalization_sweeps": thermalization, "measurement_sweeps": measurement,
              "sample_interval": sample_interval, "models": {}}
    for kind in kinds:
        L = (sizes or {}).get(kind) or (KAWAMURA["L"] if kind == "kawamura" else HE3["L"])
        print(f"[precision] {kind} L={L} R={R}: float32 chain", flush=True)
        single = chain(kind, L, R, False, thermalization, measurement, sample_interval, seed + L)
        print(f"[precision] {kind} L={L} R={R}: float64 chain", flush=True)
        double = chain(kind, L, R, True, thermalization, measurement, sample_interval, seed + L + 1000)
        entry = compare(kind, single, double, L)
        entry["L"] = L
        entry["temperatures"] = np.asarray(single["metadata"]["temperatures"]).tolist()
        entry["rate_site_updates_per_ns"] = {"float32": single["metadata"]["rate_site_updates_per_ns"],
                                             "float64": double["metadata"]["rate_site_updates_per_ns"]}
        report["models"][kind] = entry
    return report


def markdown(report):
    lines = [f"# Float32 against float64: Kawamura and He3 ({report['replicas']} replicas, "
             f"{report['thermalization_sweeps']} + {report['measurement_sweeps']} sweeps, sample every "
             f"{report['sample_interval']})", ""]
    for kind, m in report["models"].items():
        r = m["rate_site_updates_per_ns"]
        lines += [f"## {kind}, L = {m['L']}", "",
                  f"Throughput (drop-in path, exchanges and sampling included): float32 {r['float32']:.2f}, "
                  f"float64 {r['float64']:.2f} site-updates/ns. Jackknife blocks of {m['block_size_samples']} "
                  "samples (>= 4 tau_int of the slowest series).", "",
                  "| estimator | mean z^2 | max abs z | float32 at T_min | float64 at T_min |", "|---|---|---|---|---|"]
        for name, e in m["estimators"].items():
            mz = "—" if e["mean_z2"] is None else f"{e['mean_z2']:.2f}"
            xz = "—" if e["max_abs_z"] is None else f"{e['max_abs_z']:.2f}"
            lines.append(f"| {name} | {mz} | {xz} | {e['float32'][0]:.5g} ± {e['float32_err'][0]:.2g} | "
                         f"{e['float64'][0]:.5g} ± {e['float64_err'][0]:.2g} |")
        lines.append("")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--model", choices=("kawamura", "he3", "both"), default="both")
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--thermalization", type=int, default=10000)
    ap.add_argument("--measurement", type=int, default=20000)
    ap.add_argument("--sample-interval", type=int, default=10)
    ap.add_argument("--out", type=Path, default=Path("data/precision_models.json"))
    args = ap.parse_args()
    kinds = ("kawamura", "he3") if args.model == "both" else (args.model,)
    report = run(kinds, args.R, args.thermalization, args.measurement, args.sample_interval)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text(json.dumps(report))
    args.out.with_suffix(".md").write_text(markdown(report))
    print(markdown(report))


if __name__ == "__main__":
    main()

"""The three XY update paths against each other at one size, with long jackknife blocks.

XY model (J = 1, uniform window delta = 1.5), L = 64, 16 temperatures 0.8-1.0,
exchanges every 10 sweeps, three independent chains:
  * plain32: the drop-in path in float32 (per-colour kernels, the cosine of
    every angle difference, the project's float32 functions);
  * plain64: the drop-in path in float64 (libm functions);
  * fast32: the batched engine (cached cos/sin, float32 fast update).
Per temperature: the Binder cumulant U = 1 - <m^4>/(2 <m^2>^2)
(models/xy/analysis.py:47-49), the helicity modulus
[J <cx> - (J^2/T) Var(qx)] / L^2 (models/xy/analysis.py:50-53) and the energy
per site, with jackknife errors over `blocks` equal blocks of the run, and the
z-scores of each pair of paths.  Long blocks matter: the Binder cumulant and
the helicity at L = 64 decorrelate over thousands of sweeps.

    python -m main.xy_engines_check --out data/xy_engines_check.json
"""

import argparse
import json
import time
from pathlib import Path

import numpy as np


def estimators(samples, L, T, J, blocks):
    m2 = (np.asarray(samples["mx"], float)**2 + np.asarray(samples["my"], float)**2) / L**4
    cx, qx = np.asarray(samples["twist_cos_x"], float), np.asarray(samples["twist_sin_x"], float)
    e = np.asarray(samples["energy"], float) / L**2
    n = len(m2) // blocks * blocks

    def est(idx):
        M2, M4 = m2[idx].mean(0), (m2[idx]**2).mean(0)
        U = 1 - M4 / (2 * M2**2)
        Y = (J * cx[idx].mean(0) - J**2 / T * (qx[idx]**2).mean(0) + J**2 / T * qx[idx].mean(0)**2) / L**2
        return np.array([U, Y, e[idx].mean(0)])

    full = est(np.arange(n))
    groups = np.arange(n).reshape(blocks, -1)
    jk = np.array([est(np.concatenate([groups[k] for k in range(blocks) if k != b])) for b in range(blocks)])
    err = np.sqrt((blocks - 1) / blocks * ((jk - jk.mean(0))**2).sum(0))
    return full, err


def run(L=64, R=16, thermalization=20000, measurement=400000, blocks=20, J=1.0):
    from core.batched import BatchedRun, BatchedSimulation
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from core.simulation import Simulation
    from models.xy.model import XYModel
    T = np.linspace(0.8, 1.0, R)
    out = {"L": L, "temperatures": T.tolist(), "thermalization": thermalization, "measurement": measurement,
           "blocks": blocks, "paths": {}, "pairs": {}}
    res = {}
    for key, dtype, seed in (("plain32", np.float32, 11), ("plain64", np.float64, 12)):
        t0 = time.perf_counter()
        rt = Runtime(XYModel(J=J, delta=1.5, dtype=dtype, start="random"), L, T, seed=seed)
        pt = ParallelTempering(rt, exchange_interval=10, seed=seed + 100)
        s = Simulation(pt, thermalization_sweeps=thermalization, measurement_sweeps=measurement,
                       sample_interval=10).run()["samples"]
        res[key] = estimators(s, L, T, J, blocks)
        out["paths"][key] = {"seconds": time.perf_counter() - t0}
        print(f"[xy-check] {key} done", flush=True)
    t0 = time.perf_counter()
    br = BatchedRun(XYModel(J=J, delta=1.5, start="random"), L, T, seed=13, exchange_interval=10, pt_seed=113)
    s = BatchedSimulation([br], thermalization_sweeps=thermalization, measurement_sweeps=measurement,
                          sample_interval=10).run()[L]["samples"]
    res["fast32"] = estimators(s, L, T, J, blocks)
    out["paths"]["fast32"] = {"seconds": time.perf_counter() - t0}
    for key, (v, e) in res.items():
        out["paths"][key].update({"binder": v[0].tolist(), "binder_err": e[0].tolist(), "helicity": v[1].tolist(),
                                  "helicity_err": e[1].tolist(), "energy": v[2].tolist(), "energy_err": e[2].tolist()})
    for x, y in (("plain32", "plain64"), ("fast32", "plain64"), ("fast32", "plain32")):
        (a, ea), (b, eb) = res[x], res[y]
        z = (a - b) / np.sqrt(ea**2 + eb**2)
        out["pairs"][f"{x}-{y}"] = {name: {"mean_z2": float(np.mean(z[k]**2)), "max_abs_z": float(np.max(np.abs(z[k])))}
                                    for k, name in enumerate(("binder", "helicity", "energy"))}
    return out


def markdown(r):
    lines = [f"# XY update paths at L = {r['L']}: plain float32, plain float64, batched fast float32 "
             f"({len(r['temperatures'])} temperatures 0.8-1.0, {r['thermalization']} + {r['measurement']} sweeps, "
             f"{r['blocks']} jackknife blocks)", "",
             "| pair | estimator | mean z^2 | max abs z |", "|---|---|---|---|"]
    for pair, d in r["pairs"].items():
        for name, v in d.items():
            lines.append(f"| {pair} | {name} | {v['mean_z2']:.2f} | {v['max_abs_z']:.2f} |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--thermalization", type=int, default=20000)
    ap.add_argument("--measurement", type=int, default=400000)
    ap.add_argument("--blocks", type=int, default=20)
    ap.add_argument("--out", type=Path, default=Path("data/xy_engines_check.json"))
    args = ap.parse_args()
    r = run(args.L, thermalization=args.thermalization, measurement=args.measurement, blocks=args.blocks)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text(json.dumps(r))
    args.out.with_suffix(".md").write_text(markdown(r))
    print(markdown(r))


if __name__ == "__main__":
    main()

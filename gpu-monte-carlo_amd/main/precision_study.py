"""FP32 batched sweeps vs FP64 plain sweeps on the XY benchmark configuration.

Two independent chains of the same physics:
* float32: the batched engine (float32 angles and sweep arithmetic; energies
  are exact integer sums of the float32 bond terms, each rounded to a
  multiple of 2^-40);
* float64: the drop-in path (Runtime + ParallelTempering + Simulation) with
  float64 angles, betas, uniforms and sums.
Both use Gaussian-MH proposals with the same per-temperature widths (adapted
toward 0.5 acceptance on the float32 chain, then frozen) and exchanges every
10 sweeps.  For every temperature the study reports E/N, |m|, the Binder
cumulant U = 1 - <m^4>/(2<m^2>^2) (models/xy/analysis.py:47-49) and the
helicity modulus (models/xy/analysis.py:50-53), each with a blocked-jackknife
error (utils/Analysis/statistics.py:6-37), and the z-score of the float32 -
float64 difference against the combined error.

Magnetisation relaxes diffusively under local updates (tau ~ L^2 sweeps), so
jackknife blocks are sized from the measured integrated autocorrelation time
(Sokal's automatic window, c = 6): at least 4 tau_int of the slowest series,
with at least 10 blocks.  Both chains run the same dynamics (same widths and
exchange schedule), so both use the block size of the larger of their two
tau_int estimates, which are noisy at the slowest temperature; tau_int is
reported per chain.

    python -m main.precision_study --out data/precision_xy.json
"""

import argparse
import json
import time
from pathlib import Path

import numpy as np

ESTIMATORS = ("energy_per_site", "abs_m", "binder", "helicity")


def tau_int(x, c=6.0):
    """Integrated autocorrelation time of a series, in samples (automatic window)."""
    x = np.asarray(x, dtype=float)
    x = x - x.mean()
    n = len(x)
    var = float(np.dot(x, x))
    if n < 4 or var == 0.0:
        return 0.5
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / var
    tau = 0.5
    for w in range(1, n):
        tau += acf[w]
        if w >= c * tau:
            break
    return max(tau, 0.5)


def _series(samples, L):
    N = L * L
    mx, my = np.asarray(samples["mx"], float), np.asarray(samples["my"], float)
    m2 = (mx**2 + my**2) / N**2
    return {"energy_per_site": np.asarray(samples["energy"], float) / N, "abs_m": np.sqrt(m2), "m2": m2,
            "twist_cos_x": np.asarray(samples["twist_cos_x"], float),
            "twist_sin_x2": np.asarray(samples["twist_sin_x"], float)**2}


def autocorrelation(samples, L):
    """Largest tau_int over temperatures, per series."""
    out = {}
    for name, x in _series(samples, L).items():
        out[name] = float(max(tau_int(x[:, t]) for t in range(x.shape[1])))
    return out


def block_size_for(taus, n_samples, min_block=20, factor=4.0):
    b = int(np.ceil(factor * max(taus.values())))
    return int(max(min_block, min(b, n_samples // 10)))


def estimates(samples, temperatures, L, J, block_size):
    """Blocked-jackknife values and errors of the four estimators, per temperature."""
    from utils.analysis import blocked_jackknife
    N = L * L
    mx, my = np.asarray(samples["mx"], float), np.asarray(samples["my"], float)
    qx = np.asarray(samples["twist_sin_x"], float)
    m2 = (mx**2 + my**2) / N**2
    moments = {"e": np.asarray(samples["energy"], float) / N, "am": np.sqrt(m2), "m2": m2, "m4": m2**2,
               "cx": np.asarray(samples["twist_cos_x"], float), "qx": qx, "qx2": qx**2}
    T = np.asarray(temperatures, float)

    def est(m):
        helicity = (J * m["cx"] - J**2 / T * (m["qx2"] - m["qx"]**2)) / N
        return np.stack((m["e"], m["am"], 1 - m["m4"] / (2 * m["m2"]**2), helicity))

    values, errors = blocked_jackknife(moments, est, block_size)
    return {name: (values[k], errors[k]) for k, name in enumerate(ESTIMATORS)}


def run(L=128, R=32, thermalization=10000, measurement=50000, sample_interval=10, seed=170,
        adapt_sweeps=2000, exchange_interval=10, block_size=None, J=1.0, verbose=True):
    """Both chains and the comparison.  block_size None: sized from tau_int per chain."""
    from core.batched import BatchedRun, BatchedSimulation
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from core.simulation import Simulation
    from models.xy.model import XYModel

    temps = np.geomspace(0.7, 1.1, R)
    say = print if verbose else (lambda *a, **k: None)

    t0 = time.perf_counter()
    br = BatchedRun(XYModel(J=J, start="random", proposal="gaussian", sigma=1.0), L, temps, seed=seed,
                    exchange_interval=exchange_interval, pt_seed=seed + 1)
    br.adapt_sigmas(adapt_sweeps, interval=100)
    sigmas = br._sigmas_host.copy()
    run32 = BatchedSimulation([br], thermalization - adapt_sweeps, measurement, sample_interval).run()[L]
    t32 = time.perf_counter() - t0
    say(f"float32 batched chain: {t32:.1f} s", flush=True)

    t0 = time.perf_counter()
    rt = Runtime(XYModel(J=J, dtype=np.float64, start="random", proposal="gaussian", sigma=sigmas), L, temps,
                 seed=seed + 1000)
    pt = ParallelTempering(rt, exchange_interval=exchange_interval, seed=seed + 1001)
    run64 = Simulation(pt, thermalization, measurement, sample_interval).run()
    t64 = time.perf_counter() - t0
    say(f"float64 plain chain: {t64:.1f} s", flush=True)

    n_samples = len(run32["samples"]["energy"])
    tau32 = autocorrelation(run32["samples"], L)
    tau64 = autocorrelation(run64["samples"], L)
    b32 = b64 = block_size or max(block_size_for(tau32, n_samples), block_size_for(tau64, n_samples))
    e32 = estimates(run32["samples"], temps, L, J, b32)
    e64 = estimates(run64["samples"], temps, L, J, b64)
    report = {"L": L, "replicas": R, "temperatures": temps.tolist(), "sigmas": sigmas.tolist(),
              "thermalization_sweeps": thermalization, "measurement_sweeps": measurement,
              "sample_interval": sample_interval, "n_samples": n_samples,
              "block_size": {"float32": b32, "float64": b64}, "block_size_given": block_size is not None,
              "tau_int_samples": {"float32": tau32, "float64": tau64},
              "float32_seconds": t32, "float64_seconds": t64,
              "move_acceptance_float32": br.acceptance_by_temperature().tolist(), "estimators": {}}
    for name in ESTIMATORS:
        v32, s32 = e32[name]
        v64, s64 = e64[name]
        z = (v32 - v64) / np.sqrt(s32**2 + s64**2)
        report["estimators"][name] = {"float32": v32.tolist(), "float32_err": s32.tolist(),
                                      "float64": v64.tolist(), "float64_err": s64.tolist(),
                                      "z": z.tolist(), "max_abs_z": float(np.max(np.abs(z))),
                                      "chi2_per_T": float(np.mean(z**2))}
    return report


def markdown(report):
    b, tau = report["block_size"], report["tau_int_samples"]
    lines = [f"# Float32 batched vs float64 plain: XY L={report['L']}, R={report['replicas']}", "",
             f"{report['thermalization_sweeps']} thermalisation + {report['measurement_sweeps']} measurement "
             f"sweeps per chain, samples every {report['sample_interval']} sweeps ({report['n_samples']} samples). "
             "Gaussian widths adapted on the float32 chain, then frozen and shared.  Chains are independent "
             "(different seeds).", "",
             f"Jackknife blocks: {b['float32']} samples for both chains"
             + (" (given)" if report.get("block_size_given") else ", sized from the largest integrated "
                "autocorrelation time over temperatures and the two chains")
             + ". Integrated autocorrelation times (in samples):", "",
             "| series | tau_int float32 | tau_int float64 |", "|---|---|---|"]
    for k in tau["float32"]:
        lines.append(f"| {k} | {tau['float32'][k]:.1f} | {tau['float64'][k]:.1f} |")
    lines.append("")
    for name, d in report["estimators"].items():
        lines.append(f"## {name}: max |z| = {d['max_abs_z']:.2f}, mean z^2 = {d['chi2_per_T']:.2f}")
        lines.append("")
        lines.append("| T | float32 | err | float64 | err | z |")
        lines.append("|---|---|---|---|---|---|")
        for k, T in enumerate(report["temperatures"]):
            lines.append(f"| {T:.4f} | {d['float32'][k]:.6f} | {d['float32_err'][k]:.2e} | {d['float64'][k]:.6f} | "
                         f"{d['float64_err'][k]:.2e} | {d['z'][k]:+.2f} |")
        lines.append("")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--R", type=int, default=32)
    ap.add_argument("--thermalization", type=int, default=10000)
    ap.add_argument("--measurement", type=int, default=50000)
    ap.add_argument("--sample-interval", type=int, default=10)
    ap.add_argument("--block-size", type=int, default=None,
                    help="jackknife block in samples (default: from tau_int)")
    ap.add_argument("--out", type=Path, default=Path("data/precision_xy.json"))
    args = ap.parse_args()
    report = run(args.L, args.R, args.thermalization, args.measurement, args.sample_interval,
                 block_size=args.block_size)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text(json.dumps(report))
    args.out.with_suffix(".md").write_text(markdown(report))
    print("tau_int (samples):", report["tau_int_samples"], "blocks:", report["block_size"])
    for name, d in report["estimators"].items():
        print(f"{name}: max |z| {d['max_abs_z']:.2f}, mean z^2 {d['chi2_per_T']:.2f}")


if __name__ == "__main__":
    main()

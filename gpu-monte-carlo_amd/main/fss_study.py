"""Finite-size-scaling set: the packed launch against the plain path.

One packed launch (core.packed) runs Ising and XY ensembles of several sizes,
48 temperatures each, exchanges every 10 sweeps.  The same ensembles run on
the drop-in path (Runtime + ParallelTempering + Simulation, four colour
launches per sweep).  Then:

* Ising: the packed engine's update, streams and exchanges are the
  reference's, so with the same seeds its samples equal the plain path's bit
  for bit; the Binder cumulants U = 1 - <m^4>/(3 <m^2>^2)
  (models/ising/analysis.py:54-75) and their crossings for consecutive sizes
  (utils/Analysis/crossings.py:5-31) are therefore identical, and the study
  checks that they are.
* XY: the packed engine uses the float32 fast update (cached cos/sin), the
  plain path the reference's float32 update (the cosine of every angle
  difference), so the chains (independent seeds) agree in distribution.  For every size and temperature the helicity modulus and
  the Binder cumulant U = 1 - <m^4>/(2 <m^2>^2) (models/xy/analysis.py:47-54)
  are compared with blocked-jackknife errors, and so are the temperatures
  where each size's helicity crosses 2T/pi and where consecutive sizes'
  Binder cumulants cross (jackknife errors over the same blocks).

    python -m main.fss_study --out data/fss_study.json
"""

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from main.precision_study import block_size_for, tau_int

ISING_T = (2.1, 2.45)
XY_T = (0.8, 1.0)


def _ensembles(ising_sizes, xy_sizes, R, seed, xy_seed_shift=0):
    from core.batched import Ensemble
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    out = [Ensemble(IsingModel(start="random"), L, np.linspace(*ISING_T, R), seed=seed + L, exchange_interval=10,
                    pt_seed=seed + 1 + L) for L in ising_sizes]
    out += [Ensemble(XYModel(start="random"), L, np.linspace(*XY_T, R), seed=seed + L + xy_seed_shift,
                     exchange_interval=10, pt_seed=seed + 1 + L + xy_seed_shift) for L in xy_sizes]
    return out


def packed_runs(ising_sizes, xy_sizes, R, thermalization, measurement, interval, seed, chunk=None):
    from core.packed import PackedRun, PackedSimulation
    ens = _ensembles(ising_sizes, xy_sizes, R, seed)
    pr = PackedRun(ens, threads=512, task_sites=2048)
    t0 = time.perf_counter()
    runs = PackedSimulation(pr, thermalization, measurement, interval, chunk=chunk).run()
    _progress(f"packed set: {len(ens)} ensembles in {time.perf_counter() - t0:.1f} s")
    return runs, time.perf_counter() - t0


def _progress(msg):
    print(f"[fss_study] {msg}", file=sys.stderr, flush=True)


def plain_runs(ising_sizes, xy_sizes, R, thermalization, measurement, interval, seed, xy_seed_shift):
    """The same ensembles on the drop-in path; XY with shifted seeds (independent chains)."""
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from core.simulation import Simulation
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    out = {}
    t0 = time.perf_counter()
    specs = [(IsingModel(start="random"), L, ISING_T, 0) for L in ising_sizes]
    specs += [(XYModel(start="random"), L, XY_T, xy_seed_shift) for L in xy_sizes]
    for model, L, T, shift in specs:
        rt = Runtime(model, L, np.linspace(*T, R), seed=seed + L + shift)
        pt = ParallelTempering(rt, exchange_interval=10, seed=seed + 1 + L + shift)
        run = Simulation(pt, thermalization, measurement, interval).run()
        _progress(f"plain {type(model).__name__} L={L}: {time.perf_counter() - t0:.1f} s so far")
        key = L
        while key in out:
            key = (key, len(out))
        out[key] = run
    return out, time.perf_counter() - t0


def _moments(run, L, kind):
    s = run["samples"]
    N = L * L
    if kind == "ising":
        m2 = (np.asarray(s["magnetization"], float) / N)**2
        return {"m2": m2, "m4": m2**2}
    mx, my = np.asarray(s["mx"], float), np.asarray(s["my"], float)
    m2 = (mx**2 + my**2) / N**2
    qx = np.asarray(s["twist_sin_x"], float)
    return {"m2": m2, "m4": m2**2, "cx": np.asarray(s["twist_cos_x"], float), "qx": qx, "qx2": qx**2}


def _binder(m, kind):
    return 1 - m["m4"] / ((3 if kind == "ising" else 2) * m["m2"]**2)


def _helicity(m, T, L):
    return (m["cx"] - (m["qx2"] - m["qx"]**2) / T) / (L * L)


def analyse(runs, sizes, kind, T, block):
    """Per size: Binder (and XY helicity) values and errors; crossings with jackknife errors."""
    from utils.analysis import block_means, jackknife_from_blocks, linear_crossings
    keys = _keys(runs, kind, sizes)
    blocks = {L: {k: block_means(v, block) for k, v in _moments(runs[key], L, kind).items()}
              for L, key in zip(sizes, keys)}
    out = {"per_size": {}, "binder_crossings": {}, "helicity_crossings": {}}
    for L in sizes:
        v, e = jackknife_from_blocks(blocks[L], lambda m: _binder(m, kind))
        entry = {"binder": v.tolist(), "binder_err": e.tolist()}
        if kind == "xy":
            v, e = jackknife_from_blocks(blocks[L], lambda m, L=L: _helicity(m, T, L))
            entry.update(helicity=v.tolist(), helicity_err=e.tolist())
            c, ce = _crossing_jackknife({"a": blocks[L]}, lambda m, L=L: (_helicity(m["a"], T, L), 2 * T / np.pi), T)
            out["helicity_crossings"][str(L)] = {"T": c, "err": ce}
        out["per_size"][str(L)] = entry
    for L1, L2 in zip(sizes[:-1], sizes[1:]):
        c, ce = _crossing_jackknife({"a": blocks[L1], "b": blocks[L2]},
                                    lambda m: (_binder(m["a"], kind), _binder(m["b"], kind)), T)
        out["binder_crossings"][f"{L1}-{L2}"] = {"T": c, "err": ce}
    return out


def _keys(runs, kind, sizes):
    """Run keys of the ensembles of one model (Ising first, then XY, in size order)."""
    ordered = list(runs)
    ising = [k for k in ordered if runs[k]["metadata"]["model_parameters"].get("delta") is None]
    xy = [k for k in ordered if runs[k]["metadata"]["model_parameters"].get("delta") is not None]
    return ising if kind == "ising" else xy


def _crossing_jackknife(blocks, curves, T):
    """First crossing temperature of two curves of block-mean moments, with its delete-one jackknife error."""
    from utils.analysis import linear_crossings
    B = len(next(iter(next(iter(blocks.values())).values())))
    totals = {g: {k: v.sum(0) for k, v in d.items()} for g, d in blocks.items()}

    def cross(means):
        y1, y2 = curves(means)
        c = linear_crossings(T, y1, y2)
        return c[0] if c and np.isfinite(c[0]) else np.nan

    full = cross({g: {k: t / B for k, t in d.items()} for g, d in totals.items()})
    dele = np.array([cross({g: {k: (totals[g][k] - blocks[g][k][b]) / (B - 1) for k in d} for g, d in blocks.items()})
                     for b in range(B)])
    ok = np.isfinite(dele)
    if not np.isfinite(full) or ok.sum() < 2:
        return None, None
    err = float(np.sqrt((B - 1) * np.var(dele[ok])))
    return float(full), err


def run(ising_sizes=(16, 32, 64), xy_sizes=(16, 32, 64), R=48, thermalization=10000, measurement=40000,
        interval=10, seed=42, xy_seed_shift=1000, chunk=None):
    packed, t_packed = packed_runs(ising_sizes, xy_sizes, R, thermalization, measurement, interval, seed, chunk)
    plain, t_plain = plain_runs(ising_sizes, xy_sizes, R, thermalization, measurement, interval, seed,
                                xy_seed_shift)
    Ti, Tx = np.linspace(*ISING_T, R), np.linspace(*XY_T, R)
    # Ising: identical samples
    ising_equal = all(np.array_equal(packed[kp]["samples"][name], plain[kq]["samples"][name])
                      for kp, kq in zip(_keys(packed, "ising", ising_sizes), _keys(plain, "ising", ising_sizes))
                      for name in ("energy", "magnetization"))
    n = measurement // interval
    taus = {}
    for label, runs in (("packed", packed), ("plain", plain)):
        for L, key in zip(xy_sizes, _keys(runs, "xy", xy_sizes)):
            m = _moments(runs[key], L, "xy")
            taus[f"{label}_{L}"] = max(tau_int(m["m2"][:, t]) for t in range(R))
    block = block_size_for(taus, n)
    report = {"R": R, "thermalization": thermalization, "measurement": measurement, "interval": interval,
              "seconds": {"packed": t_packed, "plain": t_plain}, "ising_samples_identical": bool(ising_equal),
              "xy_block_size": block, "xy_tau_int_samples": taus}
    report["ising"] = {"packed": analyse(packed, list(ising_sizes), "ising", Ti, max(n // 20, 1)),
                       "plain": analyse(plain, list(ising_sizes), "ising", Ti, max(n // 20, 1))}
    xp, xq = analyse(packed, list(xy_sizes), "xy", Tx, block), analyse(plain, list(xy_sizes), "xy", Tx, block)
    report["xy"] = {"packed": xp, "plain": xq, "z": {}}
    for L in map(str, xy_sizes):
        for name in ("binder", "helicity"):
            a, ae = np.array(xp["per_size"][L][name]), np.array(xp["per_size"][L][name + "_err"])
            b, be = np.array(xq["per_size"][L][name]), np.array(xq["per_size"][L][name + "_err"])
            z = (a - b) / np.sqrt(ae**2 + be**2)
            report["xy"]["z"][f"{name}_{L}"] = {"mean_z2": float(np.nanmean(z**2)), "max_abs_z": float(np.nanmax(np.abs(z)))}
    for kind in ("helicity_crossings", "binder_crossings"):
        for key, c in xp[kind].items():
            d = xq[kind].get(key, {})
            if c["T"] is not None and d.get("T") is not None:
                report["xy"]["z"][f"{kind}_{key}"] = {"packed": c, "plain": d,
                                                      "z": (c["T"] - d["T"]) / np.hypot(c["err"], d["err"])}
    return report


def markdown(rep):
    lines = [f"# Finite-size-scaling set: packed launch against the plain path ({rep['R']} temperatures, "
             f"{rep['thermalization']} + {rep['measurement']} sweeps, sample every {rep['interval']})", "",
             f"Wall time: packed {rep['seconds']['packed']:.1f} s, plain {rep['seconds']['plain']:.1f} s.", "",
             f"Ising samples identical between the engines: {rep['ising_samples_identical']}.", "",
             "| Ising Binder crossing | packed T | plain T |", "|---|---|---|"]
    for key, c in rep["ising"]["packed"]["binder_crossings"].items():
        d = rep["ising"]["plain"]["binder_crossings"][key]
        fmt = lambda x: "—" if x["T"] is None else f"{x['T']:.4f} ± {x['err']:.4f}"
        lines.append(f"| {key} | {fmt(c)} | {fmt(d)} |")
    lines += ["", f"XY (jackknife blocks of {rep['xy_block_size']} samples):", "",
              "| XY comparison | packed | plain | z / mean z^2 |", "|---|---|---|---|"]
    for key, v in rep["xy"]["z"].items():
        if "packed" in v:
            lines.append(f"| {key} | {v['packed']['T']:.4f} ± {v['packed']['err']:.4f} | "
                         f"{v['plain']['T']:.4f} ± {v['plain']['err']:.4f} | {v['z']:.2f} |")
        else:
            lines.append(f"| {key} (over T) | | | {v['mean_z2']:.2f} (max abs z {v['max_abs_z']:.2f}) |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--ising-sizes", type=int, nargs="+", default=[16, 32, 64])
    ap.add_argument("--xy-sizes", type=int, nargs="+", default=[16, 32, 64])
    ap.add_argument("--R", type=int, default=48)
    ap.add_argument("--thermalization", type=int, default=10000)
    ap.add_argument("--measurement", type=int, default=40000)
    ap.add_argument("--interval", type=int, default=10)
    ap.add_argument("--chunk", type=int, default=None, help="run the packed set as graph replays of this many sweeps")
    ap.add_argument("--out", type=Path, default=Path("data/fss_study.json"))
    args = ap.parse_args()
    rep = run(tuple(args.ising_sizes), tuple(args.xy_sizes), args.R, args.thermalization, args.measurement,
              args.interval, chunk=args.chunk)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text(json.dumps(rep))
    args.out.with_suffix(".md").write_text(markdown(rep))
    print(markdown(rep))


if __name__ == "__main__":
    main()
